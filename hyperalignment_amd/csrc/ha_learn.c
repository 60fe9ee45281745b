/*
 * ha_learn.c -- learned spectral scaling of the polar iteration: spectra
 * collection, GPU training and batch inference of the 12 -> 64 -> 64 -> 2 MLP.
 *
 * The network predicts, per searchlight, y0 = log(sigma_max / ||A||_F) and
 * y1 = log10(sigma_min / sigma_max) of A = X_sl^T Y_sl from the features the
 * polar kernel computes in its prologue; the polar kernel turns them into the
 * scale of X0 = A / scale and the coefficient schedule (iteration count).
 * Labels come from the FP32 Jacobi SVD path. There is no counterpart in the
 * reference.
 */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ha_internal.h"
#include "hak.h"

#define ERR(...) fprintf(stderr, "ha_hip: " __VA_ARGS__)
#define HCK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
	ERR("%s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
	rc = kHaErrorInternal; goto fail; } } while (0)
#define KCK(x) do { int e_ = (x); if (e_ != 0) { \
	ERR("%s failed: %s (%s:%d)\n", #x, hipGetErrorString((hipError_t)e_), __FILE__, __LINE__); \
	rc = kHaErrorInternal; goto fail; } } while (0)
#define ACK(p, bytes) do { size_t b_ = (size_t)(bytes); if (b_ == 0) b_ = 16; \
	if (hipMalloc((void **)&(p), b_) != hipSuccess) { \
	ERR("device allocation of %zu bytes failed\n", b_); (void)hipGetLastError(); \
	rc = kHaErrorAlloc; goto fail; } } while (0)

int ha_hip_mlp_forward(const float *params, const float *norm, const float *X, int64_t N, float *Y) {
	int rc = ha_hip_ensure();
	if (rc) return rc;
	if (!params || !norm || !X || !Y || N < 0) return kHaErrorArg;
	hipStream_t st = (hipStream_t)ha_hip_stream();
	float *dp = NULL, *dn = NULL, *dx = NULL, *dy = NULL;
	ACK(dp, HA_HIP_MLP_NPARAM * sizeof(float));
	ACK(dn, 2 * HA_HIP_MLP_IN * sizeof(float));
	ACK(dx, (size_t)N * HA_HIP_MLP_IN * sizeof(float));
	ACK(dy, (size_t)N * HA_HIP_MLP_OUT * sizeof(float));
	HCK(hipMemcpyAsync(dp, params, HA_HIP_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(dn, norm, 2 * HA_HIP_MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(dx, X, (size_t)N * HA_HIP_MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
	KCK(hak_mlp_forward(dp, dn, dx, N, dy, st));
	HCK(hipMemcpyAsync(Y, dy, (size_t)N * HA_HIP_MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, st));
	HCK(hipStreamSynchronize(st));
fail:
	hipStreamSynchronize(st);
	if (dp) hipFree(dp);
	if (dn) hipFree(dn);
	if (dx) hipFree(dx);
	if (dy) hipFree(dy);
	return rc;
}

/* splitmix64 for the epoch shuffles */
static uint64_t sm64(uint64_t *s) {
	uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

/* Full-batch gradient (lr = 0) or Adam training over `epochs` epochs of
 * minibatches of `batch` rows in a seeded shuffled order. grad (may be NULL)
 * receives the full-batch gradient of the initial parameters; loss_hist[e]
 * (may be NULL) the mean minibatch loss of epoch e. */
static int mlp_run(float *params, const float *norm, const float *X, const float *Y, int64_t N, int32_t epochs,
                   int32_t batch, float lr, float w0, float w1, uint64_t seed, float *grad, float *loss0,
                   float *loss_hist) {
	int rc = ha_hip_ensure();
	if (rc) return rc;
	if (!params || !norm || !X || !Y || N <= 0 || batch <= 0) return kHaErrorArg;
	hipStream_t st = (hipStream_t)ha_hip_stream();
	float *dp = NULL, *dm = NULL, *dv = NULL, *dn = NULL, *dx = NULL, *dy = NULL, *part = NULL, *lp = NULL;
	float *dl = NULL, *dg = NULL;
	int32_t *didx = NULL;
	int32_t *perm = NULL;
	const int64_t maxb = batch < N ? batch : N;
	const int64_t nchunk = (N + 63) / 64; /* chunks of the largest batch (the full-batch gradient) */
	ACK(dp, HA_HIP_MLP_NPARAM * sizeof(float));
	ACK(dm, HA_HIP_MLP_NPARAM * sizeof(float));
	ACK(dv, HA_HIP_MLP_NPARAM * sizeof(float));
	ACK(dg, HA_HIP_MLP_NPARAM * sizeof(float));
	ACK(dn, 2 * HA_HIP_MLP_IN * sizeof(float));
	ACK(dx, (size_t)N * HA_HIP_MLP_IN * sizeof(float));
	ACK(dy, (size_t)N * HA_HIP_MLP_OUT * sizeof(float));
	ACK(part, (size_t)nchunk * HA_HIP_MLP_NPARAM * sizeof(float));
	ACK(lp, (size_t)nchunk * sizeof(float));
	ACK(dl, sizeof(float));
	ACK(didx, (size_t)N * sizeof(int32_t));
	HCK(hipMemcpyAsync(dp, params, HA_HIP_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(dn, norm, 2 * HA_HIP_MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(dx, X, (size_t)N * HA_HIP_MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
	HCK(hipMemcpyAsync(dy, Y, (size_t)N * HA_HIP_MLP_OUT * sizeof(float), hipMemcpyHostToDevice, st));
	KCK(hak_fill_i32((int32_t *)dm, HA_HIP_MLP_NPARAM, 0, st));
	KCK(hak_fill_i32((int32_t *)dv, HA_HIP_MLP_NPARAM, 0, st));
	if (grad || loss0) {
		KCK(hak_mlp_grad_step(dp, dm, dv, dn, dx, dy, NULL, N, w0, w1, 0.0f, 1, part, lp, dl, dg, st));
		if (grad) HCK(hipMemcpyAsync(grad, dg, HA_HIP_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, st));
		if (loss0) HCK(hipMemcpyAsync(loss0, dl, sizeof(float), hipMemcpyDeviceToHost, st));
		HCK(hipStreamSynchronize(st));
	}
	if (epochs > 0 && lr > 0) {
		perm = (int32_t *)malloc((size_t)N * sizeof(int32_t));
		if (!perm) { rc = kHaErrorAlloc; goto fail; }
		for (int64_t i = 0; i < N; ++i) perm[i] = (int32_t)i;
		uint64_t s = seed;
		int step = 0;
		for (int32_t e = 0; e < epochs; ++e) {
			for (int64_t i = N - 1; i > 0; --i) {
				int64_t j = (int64_t)(sm64(&s) % (uint64_t)(i + 1));
				int32_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
			}
			HCK(hipMemcpyAsync(didx, perm, (size_t)N * sizeof(int32_t), hipMemcpyHostToDevice, st));
			double lsum = 0;
			int nb = 0;
			for (int64_t b0 = 0; b0 < N; b0 += maxb) {
				int64_t nbat = N - b0 < maxb ? N - b0 : maxb;
				float l = 0;
				++step;
				KCK(hak_mlp_grad_step(dp, dm, dv, dn, dx, dy, didx + b0, nbat, w0, w1, lr, step, part, lp, dl,
				                      NULL, st));
				if (loss_hist) {
					HCK(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, st));
					HCK(hipStreamSynchronize(st));
					lsum += l;
					++nb;
				}
			}
			if (loss_hist) loss_hist[e] = (float)(lsum / (nb ? nb : 1));
		}
		HCK(hipMemcpyAsync(params, dp, HA_HIP_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, st));
		HCK(hipStreamSynchronize(st));
	}
fail:
	hipStreamSynchronize(st);
	free(perm);
	void *ptrs[] = {dp, dm, dv, dn, dx, dy, part, lp, dl, dg, didx};
	for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
		if (ptrs[i]) hipFree(ptrs[i]);
	return rc;
}

int ha_hip_mlp_grad(const float *params, const float *norm, const float *X, const float *Y, int64_t N, float w0,
                    float w1, float *grad, float *loss) {
	float p[HA_HIP_MLP_NPARAM];
	if (!params) return kHaErrorArg;
	memcpy(p, params, sizeof(p));
	return mlp_run(p, norm, X, Y, N, 0, (int32_t)(N < 0x7fffffff ? N : 0x7fffffff), 0.0f, w0, w1, 0, grad, loss,
	               NULL);
}

int ha_hip_mlp_train(float *params, const float *norm, const float *X, const float *Y, int64_t N, int32_t epochs,
                     int32_t batch, float lr, float w0, float w1, uint64_t seed, float *loss_hist) {
	return mlp_run(params, norm, X, Y, N, epochs, batch, lr, w0, w1, seed, NULL, NULL, loss_hist);
}
