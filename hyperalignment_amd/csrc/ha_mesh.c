/*
 * ha_mesh.c -- geodesic searchlights on a triangle mesh (host code, OpenMP).
 *
 * The reference's real-data benchmark takes its searchlights from neuroboros
 * (benchmark_neuroboros.py:153: nb.sls(lr, radius, return_dists=True)): for
 * each centre vertex, the vertices within radius mm along the cortical
 * surface and their distances, which then feed the weights
 * (ha_searchlight.c:36-92) and the searchlight passes. This builds the same
 * lists from any triangle mesh, the vertices with distance < radius (the centre
 * included, at 0), sorted by vertex id, with two distances:
 *  - along the mesh edges (Dijkstra, edge length = Euclidean distance of its
 *    end points: ha_mesh_searchlights);
 *  - geodesic across the faces (ha_mesh_searchlights_geodesic): fast marching
 *    with the planar unfolding update. When a vertex is accepted, each face
 *    around it updates its other two vertices from the edge and, where the
 *    face's third vertex is accepted as well, from the virtual point source at
 *    the two accepted distances on the far side of their edge (exact on a flat
 *    mesh whose shortest paths cross the faces, since a straight ray into a
 *    vertex enters through one edge of a face around it).
 *
 * One centre per thread: a binary heap and a distance array per thread; the
 * distances reached are reset after each centre, so a centre costs its own
 * neighbourhood only. The result does not depend on the number of threads.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "ha_internal.h"
#include "ha_mesh.h"

#define ERR(...) fprintf(stderr, "ha_mesh: " __VA_ARGS__)

typedef struct { double d; int32_t v; } HeapItem;

typedef struct {
	HeapItem *h;
	int64_t n, cap;
} Heap;

static int heap_push(Heap *hp, double d, int32_t v) {
	if (hp->n == hp->cap) {
		int64_t nc = hp->cap ? 2 * hp->cap : 256;
		HeapItem *nh = (HeapItem *)realloc(hp->h, (size_t)nc * sizeof(HeapItem));
		if (!nh) return -1;
		hp->h = nh;
		hp->cap = nc;
	}
	int64_t i = hp->n++;
	while (i > 0) {
		int64_t p = (i - 1) / 2;
		if (hp->h[p].d <= d) break;
		hp->h[i] = hp->h[p];
		i = p;
	}
	hp->h[i].d = d;
	hp->h[i].v = v;
	return 0;
}

static HeapItem heap_pop(Heap *hp) {
	HeapItem top = hp->h[0], last = hp->h[--hp->n];
	int64_t i = 0;
	for (;;) {
		int64_t c = 2 * i + 1;
		if (c >= hp->n) break;
		if (c + 1 < hp->n && hp->h[c + 1].d < hp->h[c].d) ++c;
		if (hp->h[c].d >= last.d) break;
		hp->h[i] = hp->h[c];
		i = c;
	}
	if (hp->n > 0) hp->h[i] = last;
	return top;
}

typedef struct { int32_t v; double d; } Hit;

static int cmp_hit(const void *a, const void *b) {
	const int32_t x = ((const Hit *)a)->v, y = ((const Hit *)b)->v;
	return (x > y) - (x < y);
}

static int cmp_i32(const void *a, const void *b) {
	const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
	return (x > y) - (x < y);
}

void ha_mesh_free(void *p) { free(p); }

static double dist3(const double *p, const double *q) {
	const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
	return sqrt(dx * dx + dy * dy + dz * dz);
}

/* Distance at pc from the wavefront that reached pa at da and pb at db, in the
 * plane of the face (pa, pb, pc) unfolded about its edge ab: the virtual
 * source s lies at distances da and db from a and b, on the side of ab away
 * from c; the value is |s c| when the segment s-c crosses the edge ab, else
 * infinity (the edge updates cover that case). */
static double unfold_update(const double *pa, const double *pb, const double *pc, double da, double db) {
	const double c = dist3(pa, pb), b = dist3(pa, pc), a = dist3(pb, pc);
	if (!(c > 0)) return INFINITY;
	const double cx = (b * b - a * a + c * c) / (2.0 * c), cy2 = b * b - cx * cx;
	const double sx = (da * da - db * db + c * c) / (2.0 * c), sy2 = da * da - sx * sx;
	if (!(cy2 > 0) || sy2 < 0) return INFINITY;
	const double cy = sqrt(cy2), sy = -sqrt(sy2);
	const double x0 = sx + (cx - sx) * (-sy) / (cy - sy);  /* where s-c crosses the edge's line */
	if (x0 < 0.0 || x0 > c) return INFINITY;
	const double ex = cx - sx, ey = cy - sy;
	return sqrt(ex * ex + ey * ey);
}

static int mesh_searchlights_impl(const double *coords, int32_t nv, const int32_t *faces, int32_t nf, double radius,
                                  const int32_t *centers, int32_t n_centers, int32_t **sl_indices,
                                  int32_t **sl_offsets, double **sl_dists, int geodesic);

int ha_mesh_searchlights(const double *coords, int32_t nv, const int32_t *faces, int32_t nf, double radius,
                         const int32_t *centers, int32_t n_centers, int32_t **sl_indices, int32_t **sl_offsets,
                         double **sl_dists) {
	return mesh_searchlights_impl(coords, nv, faces, nf, radius, centers, n_centers, sl_indices, sl_offsets,
	                              sl_dists, 0);
}

int ha_mesh_searchlights_geodesic(const double *coords, int32_t nv, const int32_t *faces, int32_t nf,
                                  double radius, const int32_t *centers, int32_t n_centers, int32_t **sl_indices,
                                  int32_t **sl_offsets, double **sl_dists) {
	return mesh_searchlights_impl(coords, nv, faces, nf, radius, centers, n_centers, sl_indices, sl_offsets,
	                              sl_dists, 1);
}

static int mesh_searchlights_impl(const double *coords, int32_t nv, const int32_t *faces, int32_t nf, double radius,
                                  const int32_t *centers, int32_t n_centers, int32_t **sl_indices,
                                  int32_t **sl_offsets, double **sl_dists, int geodesic) {
	if (!coords || !faces || nv <= 0 || nf <= 0 || !(radius > 0) || !isfinite(radius) || !sl_indices ||
	    !sl_offsets || !sl_dists) {
		ERR("invalid mesh searchlight arguments\n");
		return kHaErrorArg;
	}
	*sl_indices = NULL; *sl_offsets = NULL; *sl_dists = NULL;
	const int32_t nc = centers ? n_centers : nv;
	if (nc <= 0) { ERR("no centres\n"); return kHaErrorArg; }
	for (int64_t i = 0; i < 3 * (int64_t)nf; ++i)
		if (faces[i] < 0 || faces[i] >= nv) { ERR("face vertex %d out of range\n", faces[i]); return kHaErrorArg; }
	if (centers)
		for (int32_t i = 0; i < nc; ++i)
			if (centers[i] < 0 || centers[i] >= nv) { ERR("centre %d out of range\n", centers[i]); return kHaErrorArg; }

	int rc = kHaSuccess;
	int64_t *adj_ptr = (int64_t *)calloc((size_t)nv + 1, sizeof(int64_t));
	int32_t *adj = NULL;
	double *wgt = NULL;
	int64_t *vf_ptr = NULL;  /* faces around each vertex (geodesic) */
	int32_t *vf = NULL;
	Hit **lists = (Hit **)calloc((size_t)nc, sizeof(Hit *));
	int32_t *counts = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
	if (!adj_ptr || !lists || !counts) { rc = kHaErrorAlloc; goto done; }

	/* undirected edge lists per vertex (each face edge in both directions), deduplicated */
	for (int64_t f = 0; f < nf; ++f)
		for (int e = 0; e < 3; ++e) {
			adj_ptr[faces[3 * f + e] + 1] += 2;  /* both edges of the face at this vertex */
		}
	for (int32_t v = 0; v < nv; ++v) adj_ptr[v + 1] += adj_ptr[v];
	adj = (int32_t *)malloc((size_t)adj_ptr[nv] * sizeof(int32_t));
	if (!adj) { rc = kHaErrorAlloc; goto done; }
	{
		int64_t *fill = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
		if (!fill) { rc = kHaErrorAlloc; goto done; }
		memcpy(fill, adj_ptr, (size_t)nv * sizeof(int64_t));
		for (int64_t f = 0; f < nf; ++f)
			for (int e = 0; e < 3; ++e) {
				const int32_t a = faces[3 * f + e], b = faces[3 * f + (e + 1) % 3], c = faces[3 * f + (e + 2) % 3];
				adj[fill[a]++] = b;
				adj[fill[a]++] = c;
			}
		free(fill);
	}
	/* sort and deduplicate each vertex's neighbours, compacting leftward in place
	 * (the write position never passes the read position) */
	{
		int64_t *nptr = (int64_t *)malloc(((size_t)nv + 1) * sizeof(int64_t));
		if (!nptr) { rc = kHaErrorAlloc; goto done; }
		int64_t w = 0;
		for (int32_t v = 0; v < nv; ++v) {
			const int64_t b = adj_ptr[v], e = adj_ptr[v + 1];
			qsort(adj + b, (size_t)(e - b), sizeof(int32_t), cmp_i32);
			nptr[v] = w;
			for (int64_t i = b; i < e; ++i)
				if (adj[i] != v && (w == nptr[v] || adj[w - 1] != adj[i])) adj[w++] = adj[i];
		}
		nptr[nv] = w;
		memcpy(adj_ptr, nptr, ((size_t)nv + 1) * sizeof(int64_t));
		free(nptr);
	}
	wgt = (double *)malloc((size_t)(adj_ptr[nv] ? adj_ptr[nv] : 1) * sizeof(double));
	if (!wgt) { rc = kHaErrorAlloc; goto done; }
	for (int32_t v = 0; v < nv; ++v)
		for (int64_t i = adj_ptr[v]; i < adj_ptr[v + 1]; ++i) {
			const double *p = coords + 3 * (int64_t)v, *q = coords + 3 * (int64_t)adj[i];
			const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
			wgt[i] = sqrt(dx * dx + dy * dy + dz * dz);
		}

	if (geodesic) {
		vf_ptr = (int64_t *)calloc((size_t)nv + 1, sizeof(int64_t));
		vf = (int32_t *)malloc((size_t)3 * nf * sizeof(int32_t));
		int64_t *fill = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
		if (!vf_ptr || !vf || !fill) { free(fill); rc = kHaErrorAlloc; goto done; }
		for (int64_t i = 0; i < 3 * (int64_t)nf; ++i) ++vf_ptr[faces[i] + 1];
		for (int32_t v = 0; v < nv; ++v) vf_ptr[v + 1] += vf_ptr[v];
		memcpy(fill, vf_ptr, (size_t)nv * sizeof(int64_t));
		for (int64_t f = 0; f < nf; ++f)
			for (int e = 0; e < 3; ++e) vf[fill[faces[3 * f + e]]++] = (int32_t)f;
		free(fill);
	}

	int failed = 0;
#pragma omp parallel num_threads(ha_host_threads())
	{
		double *dist = (double *)malloc((size_t)nv * sizeof(double));
		uint8_t *done_v = (uint8_t *)calloc((size_t)nv, 1);
		int32_t *touched = (int32_t *)malloc((size_t)nv * sizeof(int32_t));
		Heap hp = {NULL, 0, 0};
		if (!dist || !done_v || !touched) {
#pragma omp atomic write
			failed = 1;
		} else {
			for (int32_t v = 0; v < nv; ++v) dist[v] = INFINITY;
#pragma omp for schedule(dynamic, 16)
			for (int32_t ci = 0; ci < nc; ++ci) {
				if (failed) continue;
				const int32_t c0 = centers ? centers[ci] : ci;
				int32_t nt = 0, ns = 0;
				Hit *out = NULL;
				int64_t cap = 0;
				hp.n = 0;
				dist[c0] = 0.0;
				touched[nt++] = c0;
				int bad = heap_push(&hp, 0.0, c0);
				while (!bad && hp.n > 0) {
					const HeapItem it = heap_pop(&hp);
					if (done_v[it.v] || it.d > dist[it.v]) continue;
					done_v[it.v] = 1;
					if (ns == cap) {
						cap = cap ? 2 * cap : 64;
						Hit *no = (Hit *)realloc(out, (size_t)cap * sizeof(Hit));
						if (!no) { bad = 1; break; }
						out = no;
					}
					out[ns].v = it.v;
					out[ns].d = it.d;
					++ns;
					if (!geodesic) {
						for (int64_t i = adj_ptr[it.v]; i < adj_ptr[it.v + 1]; ++i) {
							const int32_t u = adj[i];
							const double nd = it.d + wgt[i];
							if (nd < radius && nd < dist[u]) {
								if (dist[u] == INFINITY) touched[nt++] = u;
								dist[u] = nd;
								if (heap_push(&hp, nd, u)) { bad = 1; break; }
							}
						}
						continue;
					}
					/* fast marching: every face around the accepted vertex updates
					 * its two other vertices */
					const double *pv = coords + 3 * (int64_t)it.v;
					for (int64_t i = vf_ptr[it.v]; i < vf_ptr[it.v + 1] && !bad; ++i) {
						const int32_t *fc = faces + 3 * (int64_t)vf[i];
						int32_t o[2], no = 0;
						for (int e = 0; e < 3; ++e)
							if (fc[e] != it.v && no < 2) o[no++] = fc[e];
						if (no != 2) continue;  /* degenerate face */
						for (int side = 0; side < 2; ++side) {
							const int32_t u = o[side], w = o[1 - side];
							if (done_v[u]) continue;
							const double *pu = coords + 3 * (int64_t)u;
							double nd = it.d + dist3(pv, pu);
							if (done_v[w]) {
								const double t = unfold_update(pv, coords + 3 * (int64_t)w, pu, it.d, dist[w]);
								if (t < nd) nd = t;
							}
							if (nd < radius && nd < dist[u]) {
								if (dist[u] == INFINITY) touched[nt++] = u;
								dist[u] = nd;
								if (heap_push(&hp, nd, u)) { bad = 1; break; }
							}
						}
					}
				}
				for (int32_t i = 0; i < nt; ++i) { dist[touched[i]] = INFINITY; done_v[touched[i]] = 0; }
				if (bad) {
					free(out);
#pragma omp atomic write
					failed = 1;
					continue;
				}
				qsort(out, (size_t)ns, sizeof(Hit), cmp_hit);
				lists[ci] = out;
				counts[ci] = ns;
			}
		}
		free(dist); free(done_v); free(touched); free(hp.h);
	}
	if (failed) { rc = kHaErrorAlloc; goto done; }
	{
		int64_t total = 0;
		for (int32_t i = 0; i < nc; ++i) total += counts[i];
		if (total > INT32_MAX) { ERR("searchlights hold %lld entries; int32 offsets cannot\n", (long long)total); rc = kHaErrorArg; goto done; }
		int32_t *idx = (int32_t *)malloc((size_t)(total ? total : 1) * sizeof(int32_t));
		int32_t *off = (int32_t *)malloc(((size_t)nc + 1) * sizeof(int32_t));
		double *dd = (double *)malloc((size_t)(total ? total : 1) * sizeof(double));
		if (!idx || !off || !dd) { free(idx); free(off); free(dd); rc = kHaErrorAlloc; goto done; }
		int64_t o = 0;
		for (int32_t i = 0; i < nc; ++i) {
			off[i] = (int32_t)o;
			for (int32_t j = 0; j < counts[i]; ++j) { idx[o + j] = lists[i][j].v; dd[o + j] = lists[i][j].d; }
			o += counts[i];
		}
		off[nc] = (int32_t)o;
		*sl_indices = idx; *sl_offsets = off; *sl_dists = dd;
	}
done:
	if (lists) for (int32_t i = 0; i < nc; ++i) free(lists[i]);
	free(lists); free(counts); free(adj_ptr); free(adj); free(wgt); free(vf_ptr); free(vf);
	return rc;
}
