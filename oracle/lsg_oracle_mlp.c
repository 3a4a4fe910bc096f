/*
 * lsg_oracle_mlp.c -- CPU oracle of the learned-gather MLP.  TEST INFRASTRUCTURE
 * ONLY (see lsg_oracle.h).
 *
 * The surrogate network has no counterpart in the reference (it is liblsgpu's
 * addition beside the plain gather), so this file restates liblsgpu's own
 * definition: layers 32-64-64-64-64-16 (4 outputs used), ReLU between layers,
 * loss L = mean_i [ sum_{c<3} (y_ic - t_ic)^2 + lambda (y_i3 - e_i)^2 ] with
 * e_i = sqrt(mean_c (y_ic - t_ic)^2) held constant, and its gradient by
 * back-propagation, all in binary64.  With bf16=1 it rounds to bfloat16 at exactly
 * the points where the GPU kernels store BF16 (inputs, weights, post-ReLU
 * activations, the output gradient and every back-propagated pre-activation
 * gradient); with bf16=0 it is plain binary64, which tests pin by central finite
 * differences.  Parameter layout: per layer W[in][out] row-major then b[out].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int kIn[5] = {32, 64, 64, 64, 64};
static const int kOut[5] = {64, 64, 64, 64, 16};

static double bf(double x, int on)
{
    if (!on) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

uint32_t lsgo_mlp_num_params(void)
{
    uint32_t o = 0;
    for (int l = 0; l < 5; l++) o += (uint32_t)(kIn[l] * kOut[l] + kOut[l]);
    return o;
}

/* x[n x 32], t[n x 3]; y_out[n x 4] (optional), grad[nparams] (optional); returns the loss */
double lsgo_mlp_loss_grad(const float* params, const float* x, const float* t, uint32_t n, double lambda, int bf16,
                          float* y_out, double* grad)
{
    size_t wofs[5], bofs[5], o = 0;
    for (int l = 0; l < 5; l++) { wofs[l] = o; o += (size_t)kIn[l] * kOut[l]; bofs[l] = o; o += kOut[l]; }
    double* A[6]; double* Z[5];
    for (int l = 0; l < 6; l++) A[l] = (double*)malloc(sizeof(double) * n * 64);
    for (int l = 0; l < 5; l++) Z[l] = (double*)malloc(sizeof(double) * n * 64);
    for (uint32_t r = 0; r < n; r++)
        for (int k = 0; k < 32; k++) A[0][r * 32 + k] = bf(x[r * 32 + k], bf16);
    for (int l = 0; l < 5; l++) {
        const int in = kIn[l], out = kOut[l];
        for (uint32_t r = 0; r < n; r++)
            for (int c = 0; c < out; c++) {
                double s = 0;
                for (int k = 0; k < in; k++) s += A[l][r * in + k] * bf(params[wofs[l] + (size_t)k * out + c], bf16);
                double z = (double)(float)s + params[bofs[l] + c];
                if (!bf16) z = s + params[bofs[l] + c];
                Z[l][r * out + c] = z;
                A[l + 1][r * out + c] = bf(l < 4 ? (z > 0 ? z : 0) : z, bf16);
            }
    }
    double loss = 0;
    double* dz = (double*)calloc((size_t)n * 64, sizeof(double));
    double* dzn = (double*)calloc((size_t)n * 64, sizeof(double));
    for (uint32_t r = 0; r < n; r++) {
        const double* y = &Z[4][r * 16];
        if (y_out) for (int c = 0; c < 4; c++) y_out[r * 4 + c] = (float)y[c];
        double se = 0, g[4];
        for (int c = 0; c < 3; c++) { double d = y[c] - t[r * 3 + c]; se += d * d; g[c] = 2 * d / n; }
        double e = sqrt(se / 3);
        double de = y[3] - e;
        g[3] = 2 * lambda * de / n;
        loss += se + lambda * de * de;
        for (int c = 0; c < 16; c++) dz[r * 16 + c] = c < 4 ? bf(bf16 ? (double)(float)g[c] : g[c], bf16) : 0.0;
    }
    loss /= n;
    if (grad) {
        memset(grad, 0, sizeof(double) * o);
        for (int l = 4; l >= 0; l--) {
            const int in = kIn[l], out = kOut[l];
            for (int i = 0; i < in; i++)
                for (int c = 0; c < out; c++) {
                    double s = 0;
                    for (uint32_t r = 0; r < n; r++) s += A[l][r * in + i] * dz[r * out + c];
                    grad[wofs[l] + (size_t)i * out + c] = s;
                }
            for (int c = 0; c < out; c++) {
                double s = 0;
                for (uint32_t r = 0; r < n; r++) s += dz[r * out + c];
                grad[bofs[l] + c] = s;
            }
            if (l == 0) break;
            for (uint32_t r = 0; r < n; r++)
                for (int i = 0; i < in; i++) {
                    double s = 0;
                    for (int c = 0; c < out; c++) s += dz[r * out + c] * bf(params[wofs[l] + (size_t)i * out + c], bf16);
                    double g = Z[l - 1][r * in + i] > 0 ? s : 0;
                    dzn[r * in + i] = bf(bf16 ? (double)(float)g : g, bf16);
                }
            double* tmp = dz; dz = dzn; dzn = tmp;
        }
        /* padded outputs of the last layer carry no gradient */
        for (int i = 0; i < 64; i++) for (int c = 4; c < 16; c++) grad[wofs[4] + (size_t)i * 16 + c] = 0;
        for (int c = 4; c < 16; c++) grad[bofs[4] + c] = 0;
    }
    for (int l = 0; l < 6; l++) free(A[l]);
    for (int l = 0; l < 5; l++) free(Z[l]);
    free(dz); free(dzn);
    return loss;
}
