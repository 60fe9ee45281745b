/*
 * cpu_ref.c -- the CPU baseline of the bench: a C/OpenMP restatement of the
 * reference's CPU FP64 searchlight Procrustes (driver ha_searchlight.c:950-1106,
 * local solve procrustes_ws_fp64_cm at ha_searchlight.c:578-659). Test
 * infrastructure: only the bench's cpu_baseline and tests/ load it; the library
 * never does.
 *
 * Per searchlight, on one OpenMP thread with its own workspace, as the
 * reference does: the column-major gather of X[:, sl] and Y[:, sl] (memcpy of
 * whole columns), A = X_sl^T Y_sl (dgemm), the SVD of A by dgesdd ('S') with
 * dgesvd ('S', 'S') as the fallback when dgesdd fails, and T = U Vt (dgemm);
 * with reflection = 0 the LU sign of T (dgetrf) and the rank-1 flip of the
 * smallest singular pair (dger), as ha_searchlight.c:612-632. The row-weighted
 * scatter into T_out is not restated: it is k^2 additions against the k^3 of
 * the SVD, and the bench times samples of searchlights, not whole passes.
 *
 * BLAS and LAPACK come in as function pointers with the LP64 Fortran interface
 * (the bench hands over the ones scipy exports), so this file needs no BLAS
 * headers or libraries at build time.
 */
#define _GNU_SOURCE
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef void (*dgemm_fn)(char *, char *, int *, int *, int *, double *, double *, int *, double *, int *,
                         double *, double *, int *);
typedef void (*dgesdd_fn)(char *, int *, int *, double *, int *, double *, double *, int *, double *, int *,
                          double *, int *, int *, int *);
typedef void (*dgesvd_fn)(char *, char *, int *, int *, double *, int *, double *, double *, int *, double *, int *,
                          double *, int *, int *);
typedef void (*dgetrf_fn)(int *, int *, double *, int *, int *, int *);
typedef void (*dger_fn)(int *, int *, double *, double *, int *, double *, int *, double *, int *);

typedef struct {
	double *lx, *ly, *a, *backup, *s, *u, *vt, *t, *work, *det;
	int *iwork, *ipiv;
	int lwork;
} Ws;

static void ws_free(Ws *w) {
	free(w->lx); free(w->ly); free(w->a); free(w->backup); free(w->s); free(w->u); free(w->vt); free(w->t);
	free(w->work); free(w->det); free(w->iwork); free(w->ipiv);
}

static int ws_alloc(Ws *w, int nt, int kmax, dgesdd_fn dgesdd, dgesvd_fn dgesvd) {
	memset(w, 0, sizeof(*w));
	const size_t kk = (size_t)kmax * kmax;
	w->lx = malloc((size_t)nt * kmax * sizeof(double));
	w->ly = malloc((size_t)nt * kmax * sizeof(double));
	w->a = malloc(kk * sizeof(double));
	w->backup = malloc(kk * sizeof(double));
	w->s = malloc((size_t)kmax * sizeof(double));
	w->u = malloc(kk * sizeof(double));
	w->vt = malloc(kk * sizeof(double));
	w->t = malloc(kk * sizeof(double));
	w->det = malloc(kk * sizeof(double));
	w->iwork = malloc((size_t)8 * kmax * sizeof(int));
	w->ipiv = malloc((size_t)kmax * sizeof(int));
	if (!w->lx || !w->ly || !w->a || !w->backup || !w->s || !w->u || !w->vt || !w->t || !w->det || !w->iwork ||
	    !w->ipiv)
		return -1;
	/* workspace queries of both drivers at the largest size */
	int n = kmax, info = 0, q = -1;
	double opt1 = 0, opt2 = 0;
	char jz = 'S', ju = 'S', jv = 'S';
	dgesdd(&jz, &n, &n, w->a, &n, w->s, w->u, &n, w->vt, &n, &opt1, &q, w->iwork, &info);
	q = -1;
	dgesvd(&ju, &jv, &n, &n, w->a, &n, w->s, w->u, &n, w->vt, &n, &opt2, &q, &info);
	w->lwork = (int)(opt1 > opt2 ? opt1 : opt2) + 1;
	w->work = malloc((size_t)w->lwork * sizeof(double));
	return w->work ? 0 : -1;
}

/* The local solve of ha_searchlight.c:578-632 on w->lx, w->ly (nt x k each,
 * column-major); the transform lands in w->t (k x k, column-major). */
static int local_solve(Ws *w, int nt, int k, int reflection, dgemm_fn dgemm, dgesdd_fn dgesdd, dgesvd_fn dgesvd,
                       dgetrf_fn dgetrf, dger_fn dger) {
	double one = 1.0, zero = 0.0;
	char tr = 'T', no = 'N';
	int n = k, m = nt, info = 0, lw = w->lwork;
	dgemm(&tr, &no, &n, &n, &m, &one, w->lx, &m, w->ly, &m, &zero, w->a, &n);
	memcpy(w->backup, w->a, (size_t)k * k * sizeof(double));
	char jz = 'S';
	dgesdd(&jz, &n, &n, w->a, &n, w->s, w->u, &n, w->vt, &n, w->work, &lw, w->iwork, &info);
	if (info != 0) {
		memcpy(w->a, w->backup, (size_t)k * k * sizeof(double));
		char ju = 'S', jv = 'S';
		info = 0;
		dgesvd(&ju, &jv, &n, &n, w->a, &n, w->s, w->u, &n, w->vt, &n, w->work, &lw, &info);
		if (info != 0) return -2;
	}
	dgemm(&no, &no, &n, &n, &n, &one, w->u, &n, w->vt, &n, &zero, w->t, &n);
	if (!reflection && dgetrf && dger) {
		memcpy(w->det, w->t, (size_t)k * k * sizeof(double));
		int dinfo = 0;
		dgetrf(&n, &n, w->det, &n, w->ipiv, &dinfo);
		double sign = 1.0;
		for (int i = 0; i < k; ++i) {
			if (w->det[(size_t)i * k + i] < 0.0) sign = -sign;
			if (w->ipiv[i] != i + 1) sign = -sign;
		}
		if (sign < 0.0) {
			double m2 = -2.0;
			int inc = 1;
			w->s[k - 1] = -w->s[k - 1];
			dger(&n, &n, &m2, &w->u[(size_t)(k - 1) * k], &inc, &w->vt[k - 1], &n, w->t, &n);
		}
	}
	return 0;
}

/* Local transforms of the searchlights which[0 .. n_which-1] (flat lists idx,
 * off as hyperalignment_c.py:161-169) on n_threads OpenMP threads, dynamic
 * schedule, one searchlight at a time per thread. Thread t is bound to cpus[t]
 * (the caller's choice, one CPU per physical core; ncpus 0: no binding); the
 * calling thread gets its own affinity back afterwards. X, Y: nt x nv FP64,
 * column-major. T_out (may be NULL): the k x k column-major transforms one
 * after the other in the order of which. Returns 0, -1 (allocation) or -2 (an
 * SVD that both drivers failed). */
int cpu_ref_procrustes(void *gemm, void *gesdd, void *gesvd, void *getrf, void *ger, const double *X,
                       const double *Y, int nt, const int *idx, const int *off, const int *which, int n_which,
                       int reflection, int n_threads, const int *cpus, int ncpus, double *T_out) {
	dgemm_fn dgemm = (dgemm_fn)gemm;
	dgesdd_fn dgesdd = (dgesdd_fn)gesdd;
	dgesvd_fn dgesvd = (dgesvd_fn)gesvd;
	dgetrf_fn dgetrf = (dgetrf_fn)getrf;
	dger_fn dger = (dger_fn)ger;
	int kmax = 1;
	size_t *tpos = malloc((size_t)(n_which > 0 ? n_which : 1) * sizeof(size_t));
	if (!tpos) return -1;
	size_t acc = 0;
	for (int i = 0; i < n_which; ++i) {
		const int k = off[which[i] + 1] - off[which[i]];
		if (k > kmax) kmax = k;
		tpos[i] = acc;
		acc += (size_t)k * k;
	}
	int err = 0;
	cpu_set_t own;
	const int have_own = sched_getaffinity(0, sizeof(own), &own) == 0;
#pragma omp parallel num_threads(n_threads)
	{
		if (ncpus > 0) {
			cpu_set_t one;
			CPU_ZERO(&one);
			CPU_SET(cpus[omp_get_thread_num() % ncpus], &one);
			sched_setaffinity(0, sizeof(one), &one);
		}
		Ws w;
		int ok = ws_alloc(&w, nt, kmax, dgesdd, dgesvd) == 0;
		if (!ok) {
#pragma omp atomic write
			err = -1;
		}
#pragma omp for schedule(dynamic, 1)
		for (int i = 0; i < n_which; ++i) {
			if (!ok) continue;
			const int s = which[i], o = off[s], k = off[s + 1] - o;
			for (int j = 0; j < k; ++j) {
				memcpy(&w.lx[(size_t)j * nt], &X[(size_t)idx[o + j] * nt], (size_t)nt * sizeof(double));
				memcpy(&w.ly[(size_t)j * nt], &Y[(size_t)idx[o + j] * nt], (size_t)nt * sizeof(double));
			}
			const int rc = local_solve(&w, nt, k, reflection, dgemm, dgesdd, dgesvd, dgetrf, dger);
			if (rc) {
#pragma omp atomic write
				err = rc;
				continue;
			}
			if (T_out) memcpy(&T_out[tpos[i]], w.t, (size_t)k * k * sizeof(double));
		}
		ws_free(&w);
	}
	if (ncpus > 0 && have_own) sched_setaffinity(0, sizeof(own), &own);
	free(tpos);
	return err;
}
