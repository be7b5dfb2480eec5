/*
 * terrain.c -- host-side helpers of the seeded synthetic-terrain generator.
 *
 * Input preparation only (not the time step): priority-flood pit filling
 * (Barnes, Lehman & Mulla 2014, "Priority-Flood", with an epsilon gradient so
 * every filled cell drains) and D8 flow accumulation in order of decreasing
 * filled elevation.  The reference has no terrain generator; these exist to build
 * the benchmark's carved-channel DEMs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  float z;
  int64_t idx;
} node;

static void heap_push(node *h, int64_t *n, node v) {
  int64_t i = (*n)++;
  h[i] = v;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].z < h[i].z || (h[p].z == h[i].z && h[p].idx <= h[i].idx)) break;
    node t = h[p];
    h[p] = h[i];
    h[i] = t;
    i = p;
  }
}

static node heap_pop(node *h, int64_t *n) {
  node top = h[0];
  h[0] = h[--(*n)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < *n && (h[l].z < h[m].z || (h[l].z == h[m].z && h[l].idx < h[m].idx))) m = l;
    if (r < *n && (h[r].z < h[m].z || (h[r].z == h[m].z && h[r].idx < h[m].idx))) m = r;
    if (m == i) break;
    node t = h[m];
    h[m] = h[i];
    h[i] = t;
    i = m;
  }
  return top;
}

static const int DI[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
static const int DJ[8] = {-1, 0, 1, -1, 1, -1, 0, 1};

/* Fill depressions in place so every cell has a monotone path to the edge. */
int terrain_fill(int32_t ny, int32_t nx, float *z, float eps) {
  const int64_t N = (int64_t)ny * nx;
  node *heap = (node *)malloc(sizeof(node) * (size_t)N);
  uint8_t *closed = (uint8_t *)calloc((size_t)N, 1);
  if (!heap || !closed) {
    free(heap);
    free(closed);
    return -1;
  }
  int64_t hn = 0;
  for (int32_t i = 0; i < ny; ++i)
    for (int32_t j = 0; j < nx; ++j)
      if (i == 0 || j == 0 || i == ny - 1 || j == nx - 1) {
        int64_t k = (int64_t)i * nx + j;
        closed[k] = 1;
        node v = {z[k], k};
        heap_push(heap, &hn, v);
      }
  while (hn > 0) {
    node c = heap_pop(heap, &hn);
    int32_t ci = (int32_t)(c.idx / nx), cj = (int32_t)(c.idx % nx);
    for (int d = 0; d < 8; ++d) {
      int32_t ni = ci + DI[d], nj = cj + DJ[d];
      if (ni < 0 || nj < 0 || ni >= ny || nj >= nx) continue;
      int64_t k = (int64_t)ni * nx + nj;
      if (closed[k]) continue;
      closed[k] = 1;
      float floor_z = nextafterf(z[c.idx] + eps, INFINITY);
      if (z[k] < floor_z) z[k] = floor_z;
      node v = {z[k], k};
      heap_push(heap, &hn, v);
    }
  }
  free(heap);
  free(closed);
  return 0;
}

static int cmp_desc(const void *a, const void *b) {
  const node *x = (const node *)a, *y = (const node *)b;
  if (x->z > y->z) return -1;
  if (x->z < y->z) return 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* D8 flow accumulation (cells drained, including itself) on a filled DEM. */
int terrain_d8_accumulate(int32_t ny, int32_t nx, const float *z, float dx, float *acc) {
  const int64_t N = (int64_t)ny * nx;
  node *order = (node *)malloc(sizeof(node) * (size_t)N);
  if (!order) return -1;
  for (int64_t k = 0; k < N; ++k) {
    order[k].z = z[k];
    order[k].idx = k;
    acc[k] = 1.0f;
  }
  qsort(order, (size_t)N, sizeof(node), cmp_desc);
  for (int64_t o = 0; o < N; ++o) {
    int64_t k = order[o].idx;
    int32_t ci = (int32_t)(k / nx), cj = (int32_t)(k % nx);
    float best = 0.0f;
    int64_t rcv = -1;
    for (int d = 0; d < 8; ++d) {
      int32_t ni = ci + DI[d], nj = cj + DJ[d];
      if (ni < 0 || nj < 0 || ni >= ny || nj >= nx) continue;
      int64_t q = (int64_t)ni * nx + nj;
      float dist = (DI[d] != 0 && DJ[d] != 0) ? dx * 1.41421356f : dx;
      float s = (z[k] - z[q]) / dist;
      if (s > best) {
        best = s;
        rcv = q;
      }
    }
    if (rcv >= 0) acc[rcv] += acc[k];
  }
  free(order);
  return 0;
}
