// Warm-start MLP oracle -- TEST INFRASTRUCTURE ONLY (see pbf_oracle.cpp).
//
// The learned warm start has no counterpart in the reference; this file states the network the
// HIP kernels implement so their forward pass, gradients and optimiser step can be checked:
//
//   x (16) -> z1 = W1 x + b1 (64) -> h1 = bf16(act(z1))
//          -> z2 = W2 h1 + b2 (64) -> h2 = bf16(act(z2))
//          -> z3 = W3 h2 + b3 (64) -> h3 = bf16(act(z3))
//          -> y  = W4 h3 + b4 (1)
//
// act is hard-swish, z * min(max(z + 3, 0), 6) / 6, with torch's derivative at the kinks (0 up to
// and including -3, z / 3 + 1/2 between, 1 from 3 on).
//
// Weights W are row-major [out][in].  Inputs x and weights are BF16 on the device; products are
// accumulated in FP32 there and in FP64 here.  bf16() is round-to-nearest-even of the FP32
// value (applied where the device stores an activation).  With bf16_acts = 0 every rounding is
// skipped, which gives the smooth FP64 network used for the finite-difference check of the
// analytic backward.  Loss: mean squared error over the batch.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = 16, H = 64;

inline float bf16_round(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;  // inf / nan unchanged
    uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

inline double act(double z) { return z * std::min(std::max(z + 3.0, 0.0), 6.0) / 6.0; }
inline double dact(double z) { return z <= -3.0 ? 0.0 : (z < 3.0 ? z / 3.0 + 0.5 : 1.0); }

struct Params {
    const float *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4;
};

Params unpack(const float* p) {
    Params q;
    q.W1 = p;
    q.b1 = q.W1 + H * IN;
    q.W2 = q.b1 + H;
    q.b2 = q.W2 + H * H;
    q.W3 = q.b2 + H;
    q.b3 = q.W3 + H * H;
    q.W4 = q.b3 + H;
    q.b4 = q.W4 + H;
    return q;
}

}  // namespace

extern "C" {

int mlp_param_count(void) { return H * IN + H + H * H + H + H * H + H + H + 1; }

float mlp_bf16(float f) { return bf16_round(f); }

// Forward for nb samples x[nb][16] -> y[nb].  bf16_mode: round weights, inputs and stored
// activations to bf16 as the device does.
void mlp_forward(const float* params, const float* x, int nb, int bf16_mode, double* y) {
    Params P = unpack(params);
    auto w = [&](float v) -> double { return bf16_mode ? (double)bf16_round(v) : (double)v; };
    std::vector<double> a0(IN), a1(H), a2(H), a3(H);
    for (int b = 0; b < nb; ++b) {
        for (int i = 0; i < IN; ++i) a0[i] = w(x[b * IN + i]);
        for (int o = 0; o < H; ++o) {
            double z = P.b1[o];
            for (int i = 0; i < IN; ++i) z += w(P.W1[o * IN + i]) * a0[i];
            double h = act(z);
            a1[o] = bf16_mode ? (double)bf16_round((float)h) : h;
        }
        for (int o = 0; o < H; ++o) {
            double z = P.b2[o];
            for (int i = 0; i < H; ++i) z += w(P.W2[o * H + i]) * a1[i];
            double h = act(z);
            a2[o] = bf16_mode ? (double)bf16_round((float)h) : h;
        }
        for (int o = 0; o < H; ++o) {
            double z = P.b3[o];
            for (int i = 0; i < H; ++i) z += w(P.W3[o * H + i]) * a2[i];
            double h = act(z);
            a3[o] = bf16_mode ? (double)bf16_round((float)h) : h;
        }
        double yy = P.b4[0];
        for (int i = 0; i < H; ++i) yy += w(P.W4[i]) * a3[i];
        y[b] = yy;
    }
}

// Loss = mean (y - t)^2 and its gradient with respect to every parameter (same packing as
// params), in FP64.  bf16_mode as in mlp_forward; the roundings are treated as identity in the
// backward pass (the straight-through convention the device follows).
double mlp_loss_grad(const float* params, const float* x, const float* t, int nb, int bf16_mode, double* grad) {
    Params P = unpack(params);
    const int np = mlp_param_count();
    for (int i = 0; i < np; ++i) grad[i] = 0.0;
    double* gW1 = grad;
    double* gb1 = gW1 + H * IN;
    double* gW2 = gb1 + H;
    double* gb2 = gW2 + H * H;
    double* gW3 = gb2 + H;
    double* gb3 = gW3 + H * H;
    double* gW4 = gb3 + H;
    double* gb4 = gW4 + H;
    auto w = [&](float v) -> double { return bf16_mode ? (double)bf16_round(v) : (double)v; };
    std::vector<double> a0(IN), z1(H), a1(H), z2(H), a2(H), z3(H), a3(H), d3(H), d2(H), d1(H);
    double loss = 0.0;
    for (int b = 0; b < nb; ++b) {
        for (int i = 0; i < IN; ++i) a0[i] = w(x[b * IN + i]);
        for (int o = 0; o < H; ++o) {
            double z = P.b1[o];
            for (int i = 0; i < IN; ++i) z += w(P.W1[o * IN + i]) * a0[i];
            z1[o] = z;
            double h = act(z);
            a1[o] = bf16_mode ? (double)bf16_round((float)h) : h;
        }
        for (int o = 0; o < H; ++o) {
            double z = P.b2[o];
            for (int i = 0; i < H; ++i) z += w(P.W2[o * H + i]) * a1[i];
            z2[o] = z;
            double h = act(z);
            a2[o] = bf16_mode ? (double)bf16_round((float)h) : h;
        }
        for (int o = 0; o < H; ++o) {
            double z = P.b3[o];
            for (int i = 0; i < H; ++i) z += w(P.W3[o * H + i]) * a2[i];
            z3[o] = z;
            double h = act(z);
            a3[o] = bf16_mode ? (double)bf16_round((float)h) : h;
        }
        double y = P.b4[0];
        for (int i = 0; i < H; ++i) y += w(P.W4[i]) * a3[i];
        double r = y - (double)t[b];
        loss += r * r;
        double dy = 2.0 * r / nb;
        gb4[0] += dy;
        for (int i = 0; i < H; ++i) {
            gW4[i] += dy * a3[i];
            d3[i] = dy * w(P.W4[i]) * dact(z3[i]);
        }
        for (int o = 0; o < H; ++o) {
            gb3[o] += d3[o];
            for (int i = 0; i < H; ++i) gW3[o * H + i] += d3[o] * a2[i];
        }
        for (int i = 0; i < H; ++i) {
            double s = 0.0;
            for (int o = 0; o < H; ++o) s += w(P.W3[o * H + i]) * d3[o];
            d2[i] = s * dact(z2[i]);
        }
        for (int o = 0; o < H; ++o) {
            gb2[o] += d2[o];
            for (int i = 0; i < H; ++i) gW2[o * H + i] += d2[o] * a1[i];
        }
        for (int i = 0; i < H; ++i) {
            double s = 0.0;
            for (int o = 0; o < H; ++o) s += w(P.W2[o * H + i]) * d2[o];
            d1[i] = s * dact(z1[i]);
        }
        for (int o = 0; o < H; ++o) {
            gb1[o] += d1[o];
            for (int i = 0; i < IN; ++i) gW1[o * IN + i] += d1[o] * a0[i];
        }
    }
    return loss / nb;
}

// One Adam step (Kingma & Ba 2015) on FP32 parameters, the update the device optimiser applies:
// m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; p -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).
void mlp_adam(float* params, float* m, float* v, const float* grad, int n, int t, float lr, float beta1, float beta2,
              float eps) {
    double c1 = 1.0 - std::pow((double)beta1, t);
    double c2 = 1.0 - std::pow((double)beta2, t);
    for (int i = 0; i < n; ++i) {
        double g = grad[i];
        double mm = (double)beta1 * m[i] + (1.0 - beta1) * g;
        double vv = (double)beta2 * v[i] + (1.0 - beta2) * g * g;
        m[i] = (float)mm;
        v[i] = (float)vv;
        params[i] = (float)((double)params[i] - (double)lr * (mm / c1) / (std::sqrt(vv / c2) + eps));
    }
}

}  // extern "C"
