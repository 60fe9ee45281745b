/*
 * ha_sched.c -- coefficient schedules of the odd-polynomial polar iteration.
 *
 * The iteration X <- X (a I + b G + c G^2), G = X^T X, maps every singular
 * value x of X through p(x) = a x + b x^3 + c x^5. If the singular values lie
 * in [l, 1], the best such step (smallest max |1 - p(x)| on [l, 1]) is the
 * degree-5 odd minimax approximation of 1 on [l, 1], found here by the Remez
 * exchange on the four alternation points {l, r1, r2, 1} (r1, r2 are the
 * interior extrema: the roots of p'(x) = a + 3b x^2 + 5c x^4). With error E,
 * p maps [l, 1] into [1 - E, 1 + E]; dividing by 1 + E gives the next interval
 * [(1 - E)/(1 + E), 1]. This greedy composition of minimax steps is the
 * construction of Chen & Chow (2014, degree 3) and Amsel et al. (2025,
 * "The Polar Express", degree 5).
 *
 * Floor: for l -> 0 the minimax error E -> 1, i.e. the interior minimum of
 * the step, (1 - E)/(1 + E), approaches 0 and singular values sitting there
 * would be crushed to ~0 (losing their precision, or their sign to rounding).
 * Each step is therefore designed on [max(l, ELL_FLOOR), 1], where E <= 0.79
 * and the minimum image is >= 0.12; below the floor p is increasing, so the
 * new lower bound is min(p(l), (1 - E)/(1 + E)). The small-value gain
 * a / (1 + E) is ~4.1 either way.
 *
 * Safety: each step is designed on [l / (1 + SAFE), 1] and applied as
 * p(x / (1 + SAFE)), so singular values up to 1 + SAFE (from rounding, or
 * a scale slightly below sigma_max) still converge. Once 1 - l < 0.05 the
 * quintic Newton-Schulz step (15x - 10x^3 + 3x^5)/8 takes over.
 */
#include <math.h>
#include <string.h>
#include "ha_internal.h"

#define SAFE 0.01
#define ELL_FLOOR 0.03

static int solve4(double M[4][5]) {
	for (int c = 0; c < 4; ++c) {
		int piv = c;
		for (int r = c + 1; r < 4; ++r)
			if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
		if (fabs(M[piv][c]) < 1e-300) return -1;
		if (piv != c)
			for (int j = 0; j < 5; ++j) { double t = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = t; }
		for (int r = 0; r < 4; ++r) {
			if (r == c) continue;
			double f = M[r][c] / M[c][c];
			for (int j = c; j < 5; ++j) M[r][j] -= f * M[c][j];
		}
	}
	for (int r = 0; r < 4; ++r) M[r][4] /= M[r][r];
	return 0;
}

/* minimax odd quintic for 1 on [l, 1]; returns the max error, or -1 on failure */
static double remez5(double l, double *a, double *b, double *c) {
	double xs[4] = {l, l + (1 - l) * 0.25, l + (1 - l) * 0.7, 1.0};
	double A = 0, B = 0, C = 0;
	for (int it = 0; it < 60; ++it) {
		double M[4][5];
		for (int i = 0; i < 4; ++i) {
			double x = xs[i];
			M[i][0] = x; M[i][1] = x * x * x; M[i][2] = x * x * x * x * x;
			M[i][3] = (i & 1) ? -1.0 : 1.0;
			M[i][4] = 1.0;
		}
		if (solve4(M)) return -1;
		A = M[0][4]; B = M[1][4]; C = M[2][4];
		double disc = 9 * B * B - 20 * A * C;
		if (!(disc > 0) || C == 0) break;
		double y1 = (-3 * B - sqrt(disc)) / (10 * C), y2 = (-3 * B + sqrt(disc)) / (10 * C);
		if (y1 > y2) { double t = y1; y1 = y2; y2 = t; }
		double r1 = sqrt(fmin(fmax(y1, l * l), 1.0)), r2 = sqrt(fmin(fmax(y2, l * l), 1.0));
		if (!(r1 > l && r2 > r1 && r2 < 1.0)) break;
		double d = fabs(r1 - xs[1]) + fabs(r2 - xs[2]);
		xs[1] = r1; xs[2] = r2;
		if (d < 1e-15) break;
	}
	/* true max error on a grid */
	double E = 0;
	for (int i = 0; i <= 4000; ++i) {
		double x = l + (1 - l) * i / 4000.0;
		double e = fabs(1 - (A * x + B * x * x * x + C * x * x * x * x * x));
		if (e > E) E = e;
	}
	if (!(E < 1.0) || !isfinite(E)) return -1;
	*a = A; *b = B; *c = C;
	return E;
}

int ha_polar_schedule(double ell0, int max_steps, float *out, int *growth_len) {
	double l = ell0 > 0 ? ell0 : 1e-12;
	double la = 0, lb = 0, lc = 0;
	int have = 0, n = 0;
	if (growth_len) *growth_len = max_steps;
	for (int t = 0; t < max_steps; ++t) {
		double a, b, c, E;
		if (1.0 - l < 0.05) {
			if (growth_len && *growth_len == max_steps) *growth_len = t;
			a = 15.0 / 8; b = -10.0 / 8; c = 3.0 / 8;
			double x = l;
			l = (15 * x - 10 * x * x * x + 3 * x * x * x * x * x) / 8;
		} else {
			const double ld = l > ELL_FLOOR ? l : ELL_FLOOR;
			E = remez5(ld / (1 + SAFE), &a, &b, &c);
			if (E < 0) {
				if (!have) return -1;
				a = la; b = lb; c = lc;
				E = 0;
				for (int i = 0; i <= 4000; ++i) {
					double x = ld / (1 + SAFE) + (1 - ld / (1 + SAFE)) * i / 4000.0;
					double e = fabs(1 - (a * x + b * x * x * x + c * x * x * x * x * x));
					if (e > E) E = e;
				}
			}
			la = a; lb = b; lc = c; have = 1;
			double s = 1.0 / (1.0 + E);
			double sf = 1.0 / (1.0 + SAFE);
			a = a * s * sf;
			b = b * s * sf * sf * sf;
			c = c * s * sf * sf * sf * sf * sf;
			const double lo = (1 - E) / (1 + E);
			const double pl = a * l + b * l * l * l + c * l * l * l * l * l;
			l = pl < lo ? pl : lo;
		}
		out[3 * t + 0] = (float)a;
		out[3 * t + 1] = (float)b;
		out[3 * t + 2] = (float)c;
		++n;
	}
	return n;
}
