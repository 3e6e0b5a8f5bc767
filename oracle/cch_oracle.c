/* CPU oracle for the metric-dependent CCH path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with int32 weights, the reference algorithms
 * that the HIP path replaces, plus classical checkers.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
 * the checker or as the timed CPU baseline; the product never calls it.
 *
 * Arithmetic: INF = 2^31-1 absorbs additions (the IEEE inf of the reference's
 * float64, docs/technical_note.md:200); finite sums are formed in int64 and a sum
 * outside (-INF, INF) sets the overflow flag.  The reference's float thresholds
 * "< -1e-9" / "< -1e-6" (python/cch.py:274,287,664; python/cch_nb.py:570;
 * cpp/cch.cpp:534) are "< 0" on integers.
 *
 * Function -> reference it follows -> what pins it (tests/):
 *   orc_symbolic      python/cch.py:136-166 (set merge into the parent)
 *                     pinned by: cpp/cch.cpp CCH::symbolic via oracle/_ref,
 *                     closed form T = sum d(d-1)/2, chordality
 *   orc_customize     python/cch_nb.py:529-563 (serial, on-the-fly triangles)
 *                     pinned by: oracle/_ref (cpp/cch.cpp:437-463) bitwise,
 *                     restricted-path brute force (python/cch.py:9-12),
 *                     potential metamorphic test (docs/technical_note.md:80)
 *   orc_customize_par cpp/cch.cpp:468-529 (level-parallel, atomic min, OpenMP)
 *                     pinned by: equality with orc_customize
 *   orc_negcycle      python/cch_nb.py:566-572 / cpp/cch.cpp:531-536
 *                     pinned by: orc_has_negative_cycle (Bellman-Ford) verdicts,
 *                     oracle/_ref count
 *   orc_query         python/cch.py:675-732 (dist, scanned)
 *                     pinned by: Bellman-Ford, Johnson+Dijkstra, the KATs of
 *                     python/test_techniques.py:171-178, the exhaustive 3-vertex
 *                     family (363-385), oracle/_ref Query::run, scan bound
 *   orc_query_path    python/cch.py:406-507 (query_path, _unpack)
 *                     pinned by: golden vectors of the reference's own
 *                     query_path (tests/golden), path validity and length
 *   orc_update_*      python/cch.py:289-402, 666-671
 *                     pinned by: full recomputation (Theorem 4), and the closed
 *                     forms last_changed = {arcs whose full value changed} and
 *                     evaluated = |touched U tops of triangles with a changed side|
 *   orc_bellman_ford, orc_has_negative_cycle, orc_johnson, orc_dijkstra_pot
 *                     textbook algorithms (python/search.py:8-29,
 *                     python/preprocess.py:13-85), pinned against each other and
 *                     by brute force on tiny graphs
 *   orc_restricted    the closed form of a customized value: shortest v->w path
 *                     whose inner vertices rank below both ends
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define INF32 2147483647
typedef int32_t i32;
typedef int64_t i64;

static inline i32 add_inf(i32 a, i32 b, int* ovf) {
  if (a == INF32 || b == INF32) return INF32;
  i64 s = (i64)a + (i64)b;
  if (s >= (i64)INF32 || s <= -(i64)INF32) {
    *ovf = 1;
    return s > 0 ? INF32 - 1 : -INF32 + 1;
  }
  return (i32)s;
}

/* ----------------------------------------------------------- structure */

typedef struct {
  i32* v;
  i64 n, cap;
} ivec;

static void iv_push(ivec* a, i32 x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? a->cap * 2 : 4;
    a->v = (i32*)realloc(a->v, (size_t)a->cap * sizeof(i32));
  }
  a->v[a->n++] = x;
}

static int cmp_i32(const void* a, const void* b) {
  i32 x = *(const i32*)a, y = *(const i32*)b;
  return (x > y) - (x < y);
}

/* sorted-set union dst |= src \ {skip} */
static void iv_union(ivec* dst, const ivec* src, i32 skip) {
  ivec out = {0, 0, 0};
  i64 i = 0, j = 0;
  while (i < dst->n || j < src->n) {
    i32 x;
    if (j >= src->n || (i < dst->n && dst->v[i] < src->v[j])) x = dst->v[i++];
    else if (i >= dst->n || src->v[j] < dst->v[i]) x = src->v[j++];
    else { x = dst->v[i]; i++; j++; }
    if (x != skip) iv_push(&out, x);
  }
  free(dst->v);
  *dst = out;
}

/* Chordal completion in rank space by merging every vertex's upper neighbours
 * into its parent (python/cch.py:137-165).  first[n+1], parent[n] are outputs;
 * head is malloc'ed and returned through *head_out (free with orc_free).
 * Returns m+ or -1 if rank is not a permutation. */
i64 orc_symbolic(i64 n, const i64* off, const i32* tgt, const i32* rank, i64* first, i32* parent,
                 i32** head_out) {
  char* seen = (char*)calloc((size_t)n, 1);
  for (i64 v = 0; v < n; v++) {
    if (rank[v] < 0 || rank[v] >= n || seen[rank[v]]) { free(seen); return -1; }
    seen[rank[v]] = 1;
  }
  free(seen);
  ivec* up = (ivec*)calloc((size_t)n, sizeof(ivec));
  for (i64 u = 0; u < n; u++)
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      i64 v = tgt[k];
      if (v == u) continue;
      i32 ru = rank[u], rv = rank[v];
      if (ru < rv) iv_push(&up[ru], rv);
      else iv_push(&up[rv], ru);
    }
  for (i64 x = 0; x < n; x++) {
    if (up[x].n > 1) {
      qsort(up[x].v, (size_t)up[x].n, sizeof(i32), cmp_i32);
      i64 w = 1;
      for (i64 i = 1; i < up[x].n; i++)
        if (up[x].v[i] != up[x].v[w - 1]) up[x].v[w++] = up[x].v[i];
      up[x].n = w;
    }
  }
  for (i64 x = 0; x < n; x++) {
    parent[x] = -1;
    if (up[x].n == 0) continue;
    i32 p = up[x].v[0];
    parent[x] = p;
    if (up[x].n > 1) iv_union(&up[p], &up[x], p);
  }
  first[0] = 0;
  for (i64 x = 0; x < n; x++) first[x + 1] = first[x] + up[x].n;
  i32* head = (i32*)malloc((size_t)(first[n] > 0 ? first[n] : 1) * sizeof(i32));
  for (i64 x = 0; x < n; x++) {
    if (up[x].n) memcpy(head + first[x], up[x].v, (size_t)up[x].n * sizeof(i32));
    free(up[x].v);
  }
  free(up);
  *head_out = head;
  return first[n];
}

void orc_free(void* p) { free(p); }

static i64 find_arc(const i64* first, const i32* head, i64 lo, i64 hi) {
  i64 a = first[lo], b = first[lo + 1];
  while (a < b) {
    i64 mid = (a + b) / 2;
    if (head[mid] < hi) a = mid + 1;
    else b = mid;
  }
  return (a < first[lo + 1] && head[a] == hi) ? a : -1;
}

/* python/cch.py:188-204.  Returns 0, or -1 if an arc is missing. */
int orc_map_arcs(i64 n, const i64* off, const i32* tgt, const i32* rank, const i64* first,
                 const i32* head, i64* e_arc, int8_t* e_up) {
  for (i64 u = 0; u < n; u++) {
    i64 ru = rank[u];
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      i64 rv = rank[tgt[k]];
      if (ru == rv) { e_arc[k] = -1; e_up[k] = 0; continue; }
      i64 a = ru < rv ? find_arc(first, head, ru, rv) : find_arc(first, head, rv, ru);
      if (a < 0) return -1;
      e_arc[k] = a;
      e_up[k] = ru < rv ? 1 : 0;
    }
  }
  return 0;
}

i64 orc_triangles(i64 n, const i64* first) {
  i64 t = 0;
  for (i64 x = 0; x < n; x++) {
    i64 d = first[x + 1] - first[x];
    t += d * (d - 1) / 2;
  }
  return t;
}

/* ET height (cpp/cch.cpp:468-480: 1 + max level of lower neighbours over G+ arcs) */
void orc_levels(i64 n, const i64* first, const i32* head, i32* lvl) {
  for (i64 x = 0; x < n; x++) lvl[x] = 0;
  for (i64 x = 0; x < n; x++)
    for (i64 a = first[x]; a < first[x + 1]; a++) {
      i32 y = head[a];
      if (lvl[x] + 1 > lvl[y]) lvl[y] = lvl[x] + 1;
    }
}

/* --------------------------------------------------------- customization */

static void base_weights(i64 m, i64 mp, const i64* e_arc, const int8_t* e_up, const i32* w, i32* up,
                         i32* dn) {
  for (i64 a = 0; a < mp; a++) { up[a] = INF32; dn[a] = INF32; }
  for (i64 k = 0; k < m; k++) {
    i64 a = e_arc[k];
    if (a < 0) continue;
    if (e_up[k]) { if (w[k] < up[a]) up[a] = w[k]; }
    else if (w[k] < dn[a]) dn[a] = w[k];
  }
}

/* Serial customization with on-the-fly lower triangles, lowest vertex first
 * (python/cch_nb.py:529-563).  Returns 1 on int32 overflow, else 0. */
int orc_customize(i64 n, const i64* first, const i32* head, i64 m, const i64* e_arc,
                  const int8_t* e_up, const i32* w, i32* up, i32* dn) {
  const i64 mp = first[n];
  int ovf = 0;
  base_weights(m, mp, e_arc, e_up, w, up, dn);
  for (i64 x = 0; x < n; x++) {
    const i64 s = first[x], e = first[x + 1];
    for (i64 i = s; i < e; i++) {
      const i64 y = head[i];
      i64 j = first[y];
      const i64 je = first[y + 1];
      const i32 d_yx = dn[i], u_xy = up[i];
      for (i64 kk = i + 1; kk < e; kk++) {
        const i32 z = head[kk];
        while (j < je && head[j] < z) j++;
        i32 v = add_inf(d_yx, up[kk], &ovf);   /* y -> x -> z */
        if (v < up[j]) up[j] = v;
        v = add_inf(dn[kk], u_xy, &ovf);       /* z -> x -> y */
        if (v < dn[j]) dn[j] = v;
      }
    }
  }
  return ovf;
}

static inline void atomic_min_i32(i32* slot, i32 v) {
  i32 cur = __atomic_load_n(slot, __ATOMIC_RELAXED);
  while (v < cur && !__atomic_compare_exchange_n(slot, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

/* Level-parallel customization (cpp/cch.cpp:496-529): vertices of one ET height
 * run in parallel with atomic minima; levels below 512 vertices run serially. */
int orc_customize_par(i64 n, const i64* first, const i32* head, i64 m, const i64* e_arc,
                      const int8_t* e_up, const i32* w, i32* up, i32* dn, int threads) {
  const i64 mp = first[n];
  int ovf = 0;
  base_weights(m, mp, e_arc, e_up, w, up, dn);
  i32* lvl = (i32*)malloc((size_t)n * sizeof(i32));
  orc_levels(n, first, head, lvl);
  i32 maxl = 0;
  for (i64 x = 0; x < n; x++) if (lvl[x] > maxl) maxl = lvl[x];
  i64* boff = (i64*)calloc((size_t)maxl + 2, sizeof(i64));
  for (i64 x = 0; x < n; x++) boff[lvl[x] + 1]++;
  for (i32 l = 0; l <= maxl; l++) boff[l + 1] += boff[l];
  i32* bucket = (i32*)malloc((size_t)n * sizeof(i32));
  i64* pos = (i64*)malloc((size_t)(maxl + 1) * sizeof(i64));
  memcpy(pos, boff, (size_t)(maxl + 1) * sizeof(i64));
  for (i64 x = 0; x < n; x++) bucket[pos[lvl[x]]++] = (i32)x;
  for (i32 l = 0; l <= maxl; l++) {
    const i64 b0 = boff[l], b1 = boff[l + 1];
    const int par = (b1 - b0) >= 512;
#pragma omp parallel for schedule(dynamic, 64) if (par) num_threads(threads) reduction(| : ovf)
    for (i64 bi = b0; bi < b1; bi++) {
      const i64 x = bucket[bi];
      const i64 s = first[x], e = first[x + 1];
      for (i64 i = s; i < e; i++) {
        const i64 y = head[i];
        i64 j = first[y];
        const i64 je = first[y + 1];
        const i32 d_yx = dn[i], u_xy = up[i];
        for (i64 kk = i + 1; kk < e; kk++) {
          const i32 z = head[kk];
          while (j < je && head[j] < z) j++;
          atomic_min_i32(&up[j], add_inf(d_yx, up[kk], &ovf));
          atomic_min_i32(&dn[j], add_inf(dn[kk], u_xy, &ovf));
        }
      }
    }
  }
  free(lvl); free(boff); free(bucket); free(pos);
  return ovf;
}

/* python/cch_nb.py:566-572; *first_arc = smallest witness or -1 */
i64 orc_negcycle(i64 mp, const i32* up, const i32* dn, i64* first_arc) {
  i64 c = 0;
  *first_arc = -1;
  for (i64 a = 0; a < mp; a++)
    if (up[a] != INF32 && dn[a] != INF32 && (i64)up[a] + (i64)dn[a] < 0) {
      if (c == 0) *first_arc = a;
      c++;
    }
  return c;
}

/* ----------------------------------------------------------------- query */

/* python/cch.py:675-732 on ranks; df/dr/mark are caller scratch of size n, q a
 * fresh stamp (> every stamp already in mark).  Returns the distance (INF if
 * unreachable); *scanned = arcs relaxed. */
i32 orc_query(i32 rs, i32 rt, const i64* first, const i32* head, const i32* parent, const i32* up,
              const i32* dn, i32* df, i32* dr, i64* mark, i64 q, i64* scanned, int* ovf) {
  i64 x;
  for (x = rs; x >= 0; x = parent[x]) { df[x] = INF32; mark[x] = q; }
  for (x = rt; x >= 0; x = parent[x]) dr[x] = INF32;
  df[rs] = 0;
  dr[rt] = 0;
  i64 sc = 0;
  for (x = rs; x >= 0; x = parent[x]) {
    const i32 dx = df[x];
    if (dx == INF32) continue;
    sc += first[x + 1] - first[x];
    for (i64 a = first[x]; a < first[x + 1]; a++) {
      i32 v = add_inf(dx, up[a], ovf);
      if (v < df[head[a]]) df[head[a]] = v;
    }
  }
  i32 best = INF32;
  for (x = rt; x >= 0; x = parent[x]) {
    const i32 dx = dr[x];
    if (dx == INF32) continue;
    if (mark[x] == q) {
      i32 v = add_inf(dx, df[x], ovf);
      if (v < best) best = v;
    }
    sc += first[x + 1] - first[x];
    for (i64 a = first[x]; a < first[x + 1]; a++) {
      i32 v = add_inf(dx, dn[a], ovf);
      if (v < dr[head[a]]) dr[head[a]] = v;
    }
  }
  *scanned = sc;
  return best;
}

/* Many queries, optionally over OpenMP threads (each with private scratch). */
int orc_query_many(i64 n, const i64* first, const i32* head, const i32* parent, const i32* up,
                   const i32* dn, i64 nq, const i32* qs, const i32* qt, i32* dist, i64* scanned,
                   int threads) {
  int ovf = 0;
#pragma omp parallel num_threads(threads) reduction(| : ovf)
  {
    i32* df = (i32*)malloc((size_t)n * sizeof(i32));
    i32* dr = (i32*)malloc((size_t)n * sizeof(i32));
    i64* mark = (i64*)calloc((size_t)n, sizeof(i64));
    i64 stamp = 0;
#pragma omp for schedule(dynamic, 16)
    for (i64 i = 0; i < nq; i++) {
      i64 sc = 0;
      dist[i] = orc_query(qs[i], qt[i], first, head, parent, up, dn, df, dr, mark, ++stamp, &sc, &ovf);
      if (scanned) scanned[i] = sc;
    }
    free(df); free(dr); free(mark);
  }
  return ovf;
}

/* ------------------------------------------------------- partial update */

typedef struct {
  i64 n, m, mp, T;
  const i64* first;
  const i32* head;
  const i64* e_arc;
  const int8_t* e_up;
  i32 *ta, *tb, *tc;      /* lower triangles a=(x,y), b=(x,z), c=(y,z) */
  i64* coff; i32* cids;   /* by_c: triangles grouped by c (python/cch.py:308) */
  i64* soff; i32* sids;   /* as_side: triangles where an arc is a or b (309-324) */
  i32* tail;              /* 325-330 */
  i64* eoff; i32* eids;   /* arc_edges (333-347) */
  /* heap + queued set */
  i64* heap; i64 hn;
  char* queued;
} orc_upd;

/* Builds the update index (python/cch.py:289-347) for one topology.  The
 * topology arrays must stay alive while the handle is used. */
orc_upd* orc_upd_create(i64 n, const i64* first, const i32* head, i64 m, const i64* e_arc,
                        const int8_t* e_up) {
  orc_upd* u = (orc_upd*)calloc(1, sizeof(orc_upd));
  u->n = n; u->m = m; u->mp = first[n];
  u->first = first; u->head = head; u->e_arc = e_arc; u->e_up = e_up;
  i64 T = orc_triangles(n, first);
  u->T = T;
  u->ta = (i32*)malloc((size_t)(T + 1) * sizeof(i32));
  u->tb = (i32*)malloc((size_t)(T + 1) * sizeof(i32));
  u->tc = (i32*)malloc((size_t)(T + 1) * sizeof(i32));
  i64 t = 0;
  for (i64 x = 0; x < n; x++) {
    const i64 s = first[x], e = first[x + 1];
    for (i64 i = s; i < e; i++) {
      const i64 y = head[i];
      i64 j = first[y];
      for (i64 k = i + 1; k < e; k++) {
        while (head[j] < head[k]) j++;
        u->ta[t] = (i32)i; u->tb[t] = (i32)k; u->tc[t] = (i32)j;
        t++;
      }
    }
  }
  const i64 mp = u->mp;
  u->coff = (i64*)calloc((size_t)mp + 1, sizeof(i64));
  u->cids = (i32*)malloc((size_t)(T + 1) * sizeof(i32));
  for (i64 i = 0; i < T; i++) u->coff[u->tc[i] + 1]++;
  for (i64 a = 0; a < mp; a++) u->coff[a + 1] += u->coff[a];
  i64* pos = (i64*)malloc((size_t)(mp + 1) * sizeof(i64));
  memcpy(pos, u->coff, (size_t)mp * sizeof(i64));
  for (i64 i = 0; i < T; i++) u->cids[pos[u->tc[i]]++] = (i32)i;
  u->soff = (i64*)calloc((size_t)mp + 1, sizeof(i64));
  u->sids = (i32*)malloc((size_t)(2 * T + 1) * sizeof(i32));
  for (i64 i = 0; i < T; i++) { u->soff[u->ta[i] + 1]++; u->soff[u->tb[i] + 1]++; }
  for (i64 a = 0; a < mp; a++) u->soff[a + 1] += u->soff[a];
  memcpy(pos, u->soff, (size_t)mp * sizeof(i64));
  for (i64 i = 0; i < T; i++) u->sids[pos[u->ta[i]]++] = (i32)i;
  for (i64 i = 0; i < T; i++) u->sids[pos[u->tb[i]]++] = (i32)i;
  u->tail = (i32*)malloc((size_t)(mp + 1) * sizeof(i32));
  for (i64 x = 0; x < n; x++)
    for (i64 a = first[x]; a < first[x + 1]; a++) u->tail[a] = (i32)x;
  u->eoff = (i64*)calloc((size_t)mp + 1, sizeof(i64));
  for (i64 k = 0; k < m; k++) if (e_arc[k] >= 0) u->eoff[e_arc[k] + 1]++;
  for (i64 a = 0; a < mp; a++) u->eoff[a + 1] += u->eoff[a];
  u->eids = (i32*)malloc((size_t)(u->eoff[mp] + 1) * sizeof(i32));
  memcpy(pos, u->eoff, (size_t)mp * sizeof(i64));
  for (i64 k = 0; k < m; k++) if (e_arc[k] >= 0) u->eids[pos[e_arc[k]]++] = (i32)k;
  free(pos);
  u->heap = (i64*)malloc((size_t)(mp + 1) * sizeof(i64));
  u->queued = (char*)calloc((size_t)mp + 1, 1);
  return u;
}

void orc_upd_destroy(orc_upd* u) {
  if (!u) return;
  free(u->ta); free(u->tb); free(u->tc); free(u->coff); free(u->cids); free(u->soff);
  free(u->sids); free(u->tail); free(u->eoff); free(u->eids); free(u->heap); free(u->queued);
  free(u);
}

/* heap of arcs keyed by (tail rank, arc id) -- the tuple order of python/cch.py:376 */
static inline int key_less(const orc_upd* u, i64 a, i64 b) {
  return u->tail[a] < u->tail[b] || (u->tail[a] == u->tail[b] && a < b);
}
static void hpush(orc_upd* u, i64 a) {
  i64 i = u->hn++;
  u->heap[i] = a;
  while (i > 0) {
    i64 p = (i - 1) / 2;
    if (!key_less(u, u->heap[i], u->heap[p])) break;
    i64 t = u->heap[i]; u->heap[i] = u->heap[p]; u->heap[p] = t;
    i = p;
  }
}
static i64 hpop(orc_upd* u) {
  i64 top = u->heap[0];
  u->heap[0] = u->heap[--u->hn];
  i64 i = 0;
  while (1) {
    i64 l = 2 * i + 1, r = l + 1, s = i;
    if (l < u->hn && key_less(u, u->heap[l], u->heap[s])) s = l;
    if (r < u->hn && key_less(u, u->heap[r], u->heap[s])) s = r;
    if (s == i) break;
    i64 t = u->heap[i]; u->heap[i] = u->heap[s]; u->heap[s] = t;
    i = s;
  }
  return top;
}

/* python/cch.py:349-402.  w[m] is updated in place with the changes; up/dn and
 * base_up/base_dn are the current customization and base values.  changed[]
 * (capacity mp) receives last_changed in pop order.  Returns evaluated;
 * *n_changed, *ovf are outputs. */
i64 orc_update(orc_upd* u, i32* w, i32* up, i32* dn, i32* bu, i32* bd, i64 k, const i64* edge_idx,
               const i32* new_w, i32* changed, i64* n_changed, int* ovf) {
  u->hn = 0;
  for (i64 i = 0; i < k; i++) w[edge_idx[i]] = new_w[i];
  i64 nt = 0;
  i64* touched = (i64*)malloc((size_t)(k + 1) * sizeof(i64));
  for (i64 i = 0; i < k; i++) {
    i64 a = u->e_arc[edge_idx[i]];
    if (a >= 0 && !u->queued[a]) { u->queued[a] = 1; touched[nt++] = a; }
  }
  for (i64 i = 0; i < nt; i++) {
    i64 a = touched[i];
    i32 nu = INF32, nd = INF32;
    for (i64 j = u->eoff[a]; j < u->eoff[a + 1]; j++) {
      i32 e = u->eids[j];
      if (u->e_up[e]) { if (w[e] < nu) nu = w[e]; }
      else if (w[e] < nd) nd = w[e];
    }
    bu[a] = nu; bd[a] = nd;
    hpush(u, a);
  }
  i64 evaluated = 0, nc = 0;
  i64* qlist = (i64*)malloc((size_t)(u->mp + 1) * sizeof(i64));
  i64 nq = 0;
  for (i64 i = 0; i < nt; i++) qlist[nq++] = touched[i];
  while (u->hn > 0) {
    i64 c = hpop(u);
    evaluated++;
    i32 nu = bu[c], nd = bd[c];
    for (i64 j = u->coff[c]; j < u->coff[c + 1]; j++) {
      i32 t = u->cids[j];
      i32 a = u->ta[t], b = u->tb[t];
      i32 v = add_inf(dn[a], up[b], ovf);
      if (v < nu) nu = v;
      v = add_inf(dn[b], up[a], ovf);
      if (v < nd) nd = v;
    }
    if (nu != up[c] || nd != dn[c]) {
      up[c] = nu; dn[c] = nd;
      changed[nc++] = (i32)c;
      for (i64 j = u->soff[c]; j < u->soff[c + 1]; j++) {
        i32 c2 = u->tc[u->sids[j]];
        if (!u->queued[c2]) { u->queued[c2] = 1; qlist[nq++] = c2; hpush(u, c2); }
      }
    }
  }
  for (i64 i = 0; i < nq; i++) u->queued[qlist[i]] = 0;
  free(qlist);
  free(touched);
  *n_changed = nc;
  return evaluated;
}

/* python/cch.py:234-247 base values (for seeding orc_update's base arrays) */
void orc_base(i64 m, i64 mp, const i64* e_arc, const int8_t* e_up, const i32* w, i32* bu, i32* bd) {
  base_weights(m, mp, e_arc, e_up, w, bu, bd);
}

/* ----------------------------------------------------- classical checkers */

/* Bellman-Ford from s over the directed CSR (python/search.py:8-29): dist[n]
 * (INF unreachable).  Returns 1 if a negative cycle is reachable from s. */
int orc_bellman_ford(i64 n, const i64* off, const i32* tgt, const i32* w, i32 s, i64* dist) {
  const i64 INF64 = INT64_MAX / 4;
  for (i64 v = 0; v < n; v++) dist[v] = INF64;
  dist[s] = 0;
  for (i64 it = 0; it < n; it++) {
    int changed = 0;
    for (i64 u = 0; u < n; u++) {
      if (dist[u] >= INF64) continue;
      for (i64 k = off[u]; k < off[u + 1]; k++) {
        if (w[k] == INF32) continue;
        i64 nd = dist[u] + w[k];
        if (nd < dist[tgt[k]]) { dist[tgt[k]] = nd; changed = 1; }
      }
    }
    if (!changed) {
      for (i64 v = 0; v < n; v++) if (dist[v] >= INF64) dist[v] = INF32;
      return 0;
    }
  }
  for (i64 v = 0; v < n; v++) if (dist[v] >= INF64) dist[v] = INF32;
  return 1;
}

/* Negative-cycle test of the whole graph: Bellman-Ford from a virtual source
 * (python/test_techniques.py:347-358). */
int orc_has_negative_cycle(i64 n, const i64* off, const i32* tgt, const i32* w) {
  i64* d = (i64*)calloc((size_t)n, sizeof(i64));
  for (i64 it = 0; it <= n; it++) {
    int changed = 0;
    for (i64 u = 0; u < n; u++)
      for (i64 k = off[u]; k < off[u + 1]; k++) {
        if (w[k] == INF32) continue;
        i64 nd = d[u] + w[k];
        if (nd < d[tgt[k]]) { d[tgt[k]] = nd; changed = 1; }
      }
    if (!changed) { free(d); return 0; }
  }
  free(d);
  return 1;
}

/* Johnson potential by FIFO Bellman-Ford from a virtual source
 * (python/preprocess.py:13-39, python/cch_nb.py:647-673): p[n] <= 0.
 * Returns 0, or 1 if a negative cycle exists. */
int orc_johnson(i64 n, const i64* off, const i32* tgt, const i32* w, i64* p) {
  i64* q = (i64*)malloc((size_t)(n + 1) * sizeof(i64));
  char* inq = (char*)malloc((size_t)n);
  i64* cnt = (i64*)calloc((size_t)n, sizeof(i64));
  for (i64 v = 0; v < n; v++) { p[v] = 0; q[v] = v; inq[v] = 1; }
  i64 h = 0, t = n % (n + 1), size = n, cap = n + 1;
  int neg = 0;
  while (size > 0 && !neg) {
    i64 u = q[h]; h = (h + 1) % cap; size--; inq[u] = 0;
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      if (w[k] == INF32) continue;
      i64 v = tgt[k], nd = p[u] + w[k];
      if (nd < p[v]) {
        p[v] = nd;
        if (++cnt[v] > n) { neg = 1; break; }
        if (!inq[v]) { inq[v] = 1; q[t] = v; t = (t + 1) % cap; size++; }
      }
    }
  }
  free(q); free(inq); free(cnt);
  return neg;
}

/* Dijkstra on reduced weights w + p[u] - p[v] >= 0 (python/cch_nb.py:618-644),
 * binary heap.  Returns dist(s, t) in original weights, INF32 if unreachable. */
i64 orc_dijkstra_pot(i64 n, const i64* off, const i32* tgt, const i32* w, const i64* p, i32 s, i32 t,
                     i64* dist, char* done, i64* heap_k, i32* heap_v) {
  const i64 INF64 = INT64_MAX / 4;
  for (i64 v = 0; v < n; v++) { dist[v] = INF64; done[v] = 0; }
  i64 hn = 0;
  dist[s] = 0;
  heap_k[hn] = 0; heap_v[hn] = s; hn++;
  while (hn > 0) {
    i64 d = heap_k[0];
    i32 u = heap_v[0];
    hn--;
    heap_k[0] = heap_k[hn]; heap_v[0] = heap_v[hn];
    for (i64 i = 0;;) {
      i64 l = 2 * i + 1, r = l + 1, sm = i;
      if (l < hn && heap_k[l] < heap_k[sm]) sm = l;
      if (r < hn && heap_k[r] < heap_k[sm]) sm = r;
      if (sm == i) break;
      i64 tk = heap_k[i]; heap_k[i] = heap_k[sm]; heap_k[sm] = tk;
      i32 tv = heap_v[i]; heap_v[i] = heap_v[sm]; heap_v[sm] = tv;
      i = sm;
    }
    if (done[u]) continue;
    done[u] = 1;
    if (u == t) break;
    for (i64 k = off[u]; k < off[u + 1]; k++) {
      if (w[k] == INF32) continue;
      i32 v = tgt[k];
      i64 rw = (i64)w[k] + p[u] - p[v];
      if (rw < 0) rw = 0;
      i64 nd = d + rw;
      if (nd < dist[v]) {
        dist[v] = nd;
        i64 i = hn++;
        heap_k[i] = nd; heap_v[i] = v;
        while (i > 0) {
          i64 pa = (i - 1) / 2;
          if (heap_k[pa] <= heap_k[i]) break;
          i64 tk = heap_k[i]; heap_k[i] = heap_k[pa]; heap_k[pa] = tk;
          i32 tv = heap_v[i]; heap_v[i] = heap_v[pa]; heap_v[pa] = tv;
          i = pa;
        }
      }
    }
  }
  if (dist[t] >= INF64) return INF32;
  return dist[t] - p[s] + p[t];
}

/* Closed form of one customized value (python/cch.py:9-12): the shortest v -> w
 * walk whose inner vertices all rank below min(rank v, rank w), by Bellman-Ford
 * restricted to those vertices.  Vertices are ORIGINAL ids.  INF32 if none. */
i64 orc_restricted(i64 n, const i64* off, const i32* tgt, const i32* w, const i32* rank, i32 v,
                   i32 wv, i64* dist) {
  const i64 INF64 = INT64_MAX / 4;
  const i32 lim = rank[v] < rank[wv] ? rank[v] : rank[wv];
  for (i64 x = 0; x < n; x++) dist[x] = INF64;
  dist[v] = 0;
  i64 best = INF64;
  for (i64 it = 0; it <= n; it++) {
    int changed = 0;
    for (i64 u = 0; u < n; u++) {
      if (dist[u] >= INF64) continue;
      if (u != v && rank[u] >= lim) continue; /* only v and low vertices expand */
      for (i64 k = off[u]; k < off[u + 1]; k++) {
        if (w[k] == INF32) continue;
        i32 x = tgt[k];
        i64 nd = dist[u] + w[k];
        if (x == wv) { if (nd < best) best = nd; continue; }
        if (x == v || rank[x] >= lim) continue;
        if (nd < dist[x]) { dist[x] = nd; changed = 1; }
      }
    }
    if (!changed) break;
  }
  return best >= INF64 ? INF32 : best;
}

/* Customization with stop_on_negative_cycle (python/cch.py:264-278): triangles
 * in (x, i, k) order; before relaxing a triangle whose first arc is i, the arcs
 * of every not yet checked vertex whose row ends at or before i are tested.
 * Returns the first witness arc found that way, or -1 if the loop finishes
 * without one (the reference then returns the full list). */
i64 orc_customize_stop(i64 n, const i64* first, const i32* head, i64 m, const i64* e_arc,
                       const int8_t* e_up, const i32* w, i32* up, i32* dn) {
  const i64 mp = first[n];
  int ovf = 0;
  base_weights(m, mp, e_arc, e_up, w, up, dn);
  i64 xc = 0;
  for (i64 x = 0; x < n; x++) {
    const i64 s = first[x], e = first[x + 1];
    for (i64 i = s; i < e; i++) {
      const i64 y = head[i];
      i64 j = first[y];
      const i64 je = first[y + 1];
      for (i64 kk = i + 1; kk < e; kk++) {
        while (xc < n && first[xc + 1] <= i) {
          for (i64 aa = first[xc]; aa < first[xc + 1]; aa++)
            if (up[aa] != INF32 && dn[aa] != INF32 && (i64)up[aa] + (i64)dn[aa] < 0) return aa;
          xc++;
        }
        const i32 z = head[kk];
        while (j < je && head[j] < z) j++;
        i32 v = add_inf(dn[i], up[kk], &ovf);
        if (v < up[j]) up[j] = v;
        v = add_inf(dn[kk], up[i], &ovf);
        if (v < dn[j]) dn[j] = v;
      }
    }
  }
  return -1;
}

/* ---------------------------------------------------------- path unpacking */

/* Shortest path with its vertices (python/cch.py:406-507, query_path, _unpack,
 * _lower_triangle_vertices), ranks in and out.  lfirst/lnbr: the lower
 * neighbours of every vertex, ascending (the x of the lower triangles of an
 * arc (y, z) in the order of the reference's by_c groups).  tail[a]: lower end
 * of arc a.  Writes up to cap ranks to path; *len = path length (-1 if
 * unreachable, -2 if the unpacking passed 4n + 16 vertices, which only a
 * metric with a negative cycle makes it do: it stops there).  Returns the
 * distance (INF32 if unreachable). */
i64 orc_query_path(i64 n, const i64* first, const i32* head, const i32* parent, const i32* tail,
                   const i64* lfirst, const i32* lnbr, const i32* up, const i32* dn, i32 rs, i32 rt,
                   i32* path, i64 cap, i64* len) {
  const i64 INF64 = INF32;
  i64* df = (i64*)malloc((size_t)n * sizeof(i64));
  i64* dr = (i64*)malloc((size_t)n * sizeof(i64));
  i64* pf = (i64*)malloc((size_t)n * sizeof(i64));
  i64* pr = (i64*)malloc((size_t)n * sizeof(i64));
  char* anc = (char*)calloc((size_t)n, 1);
  for (i64 v = 0; v < n; v++) { df[v] = dr[v] = INF64; pf[v] = pr[v] = -1; }
  df[rs] = 0;
  dr[rt] = 0;
  for (i64 x = rs; x >= 0; x = parent[x]) {
    if (df[x] >= INF64) continue;
    for (i64 a = first[x]; a < first[x + 1]; a++) {
      if (up[a] == INF32) continue;
      i64 v = df[x] + up[a];
      if (v < df[head[a]]) { df[head[a]] = v; pf[head[a]] = a; }
    }
  }
  for (i64 x = rt; x >= 0; x = parent[x]) {
    if (dr[x] >= INF64) continue;
    for (i64 a = first[x]; a < first[x + 1]; a++) {
      if (dn[a] == INF32) continue;
      i64 v = dr[x] + dn[a];
      if (v < dr[head[a]]) { dr[head[a]] = v; pr[head[a]] = a; }
    }
  }
  for (i64 x = rt; x >= 0; x = parent[x]) anc[x] = 1;
  i64 best = INF64, peak = -1;
  for (i64 x = rs; x >= 0; x = parent[x])
    if (anc[x] && df[x] < INF64 && dr[x] < INF64 && df[x] + dr[x] < best) { best = df[x] + dr[x]; peak = x; }
  *len = -1;
  if (peak >= 0) {
    /* legs as (from, to) rank pairs: up from rs to the peak, then down to rt */
    i64 nl = 0, cl = 64;
    i32* legs = (i32*)malloc((size_t)cl * 2 * sizeof(i32));
    for (i64 x = peak; x != rs;) {
      i32 v = tail[pf[x]];
      if (nl == cl) { cl *= 2; legs = (i32*)realloc(legs, (size_t)cl * 2 * sizeof(i32)); }
      legs[2 * nl] = v; legs[2 * nl + 1] = (i32)x; nl++;
      x = v;
    }
    for (i64 i = 0, j = nl - 1; i < j; i++, j--) {  /* reverse */
      i32 a0 = legs[2 * i], a1 = legs[2 * i + 1];
      legs[2 * i] = legs[2 * j]; legs[2 * i + 1] = legs[2 * j + 1];
      legs[2 * j] = a0; legs[2 * j + 1] = a1;
    }
    for (i64 x = peak; x != rt;) {
      i32 v = tail[pr[x]];
      if (nl == cl) { cl *= 2; legs = (i32*)realloc(legs, (size_t)cl * 2 * sizeof(i32)); }
      legs[2 * nl] = (i32)x; legs[2 * nl + 1] = v; nl++;
      x = v;
    }
    i64 L = 0;
    if (L < cap) path[L] = nl ? legs[0] : rs;
    L++;
    i64 sc = 64, sp = 0;
    const i64 lim = 4 * n + 16;
    i32* st = (i32*)malloc((size_t)sc * 2 * sizeof(i32));
    for (i64 l = 0; l < nl && L <= lim; l++) {
      st[0] = legs[2 * l]; st[1] = legs[2 * l + 1]; sp = 1;
      while (sp > 0 && L <= lim) {
        sp--;
        const i32 v = st[2 * sp], w = st[2 * sp + 1];
        const i32 lo = v < w ? v : w, hi = v < w ? w : v;
        const i64 a = find_arc(first, head, lo, hi);
        const i64 cost = v < w ? up[a] : dn[a];
        i32 mid = -1;
        /* lower triangles of arc (lo, hi): x < lo adjacent to both, ascending */
        for (i64 k = lfirst[lo]; k < lfirst[lo + 1] && mid < 0; k++) {
          const i32 u = lnbr[k];
          const i64 av = find_arc(first, head, u, v), aw = find_arc(first, head, u, w);
          if (av < 0 || aw < 0) continue;
          const i64 c1 = dn[av], c2 = up[aw];
          if (c1 != INF32 && c2 != INF32 && cost != INF32 && c1 + c2 == cost) mid = u;
        }
        if (mid < 0) {
          if (L < cap) path[L] = w;
          L++;
        } else {
          if (sp + 2 > sc) { sc *= 2; st = (i32*)realloc(st, (size_t)sc * 2 * sizeof(i32)); }
          st[2 * sp] = mid; st[2 * sp + 1] = w; sp++;   /* popped second */
          st[2 * sp] = v; st[2 * sp + 1] = mid; sp++;   /* popped first */
        }
      }
    }
    free(st);
    free(legs);
    *len = L > lim ? -2 : L;
  }
  free(df); free(dr); free(pf); free(pr); free(anc);
  return peak >= 0 ? best : INF32;
}
