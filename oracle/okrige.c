/* CPU restatement of the reference kriging hot path -- TEST INFRASTRUCTURE.
 *
 * Parity oracle for the HIP product library.  Only tests/, the smoke entry
 * point and bench.py's cpu_baseline leg load it.  Built with
 * -ffp-contract=off so squared distances and table interpolation round as
 * in the reference's x86-64 build (no fused multiply-add).
 *
 * Restated from (paths relative to the reference repository root):
 *   kd-tree          src/libkriging/kdtree2_maxidx.f90:307-461 (heap),
 *                    :719-998 (build), :1043-1084,1383-1613 (query),
 *                    :1860-1914 (heapsort)
 *   search           src/libkriging/kriging.F90:490-666, :1093-1154 (sector)
 *   filters/sort     src/libkriging/kriging_base.F90:2491-2505, :3510-3524
 *   covariance       src/libkriging/variogram.f90:185-382, 497-771, 853-869,
 *                    src/libkriging/vgmfunc.f90:48-100,
 *                    src/libkriging/rotation.f90:51-95
 *   assembly         src/libkriging/kriging.F90:714-849 (LHS), 864-939 (RHS)
 *   solve            src/libkriging/solver.f90:224-300,
 *                    src/libkriging/kriging_base.F90:3244-3398 (fallbacks)
 *   variance         src/libkriging/kriging.F90:968-1028
 *   estimate         src/libkriging/kriging_base.F90:3407-3503
 */
#include "okrige.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXS 16

/* ------------------------------------------------------------------------ */
/* constants: src/libkriging/common.f90:7-12                                 */
/* ------------------------------------------------------------------------ */
static double PI_F(void) { return 4.0 * atan(1.0); }
#define EPSLON 1.e-10

/* ------------------------------------------------------------------------ */
/* correlation shapes: src/libkriging/vgmfunc.f90:48-87                      */
/* ------------------------------------------------------------------------ */
static double corefunc(int t, double r) {
  const double pi = PI_F();
  switch (t) {
    case 0: return 0.0;
    case 1: return (r < 1.0) ? (1.0 - 1.5 * r) + 0.5 * (r * r * r) : 0.0;
    case 2: return exp(-3.0 * r);
    case 3: return exp(-3.0625 * (r * r));
    case 4: return cos(pi * r);
    case 5: return (r < 1.0) ? 1.0 - pow(r, 1.5) : 0.0;
    case 6: { double a = 1.0 - r * r; return (r < 1.0) ? a * a : 0.0; }
    case 7: return (r < 1.0) ? 1.0 - (2.0 * r * sqrt(1.0 - r * r) + 2.0 * asin(r)) / pi : 0.0;
    case 8: return (r < 1.0) ? 1.0 - r : 0.0;
    case 9: { double s = sin(pi * r); return exp(-2.0 * (s * s)); }
    case 10: return exp(-3.0 * r) * cos(pi * r);
    default: return 0.0;
  }
}
/* vgmfunc.f90:91-100 */
static int has_tail(int t) { return t == 2 || t == 3 || t == 4 || t == 9 || t == 10; }

/* ------------------------------------------------------------------------ */
/* rotation: src/libkriging/rotation.f90:51-95 (3x3, column-major m[r+3c])   */
/* ------------------------------------------------------------------------ */
static void matmul3(const double *a, const double *b, double *c) {
  double t[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + a[i + 3 * k] * b[k + 3 * j];
      t[i + 3 * j] = s;
    }
  memcpy(c, t, sizeof t);
}

static void calc_rotmat(double ang1, double ang2, double ang3, double anis1, double anis2,
                        double *R) {
  const double d2r = PI_F() / 180.0;
  memset(R, 0, 9 * sizeof(double));
  R[0] = 1.0 / anis1;
  R[4] = 1.0;
  R[8] = 1.0 / anis2;
  double T[9];
  if (fabs(ang3) > EPSLON) { /* about Y */
    double a = -ang3 * d2r, s = sin(a), c = cos(a);
    double t[9] = {c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c};
    memcpy(T, t, sizeof t);
    matmul3(R, T, R);
  }
  if (fabs(ang2) > EPSLON) { /* about X */
    double a = -ang2 * d2r, s = sin(a), c = cos(a);
    double t[9] = {1.0, 0.0, 0.0, 0.0, c, s, 0.0, -s, c};
    memcpy(T, t, sizeof t);
    matmul3(R, T, R);
  }
  if (fabs(ang1) > EPSLON) { /* about Z */
    double a = ang1 * d2r, s = sin(a), c = cos(a);
    double t[9] = {c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0};
    memcpy(T, t, sizeof t);
    matmul3(R, T, R);
  }
}

/* ------------------------------------------------------------------------ */
/* covariance model (one vgm_struct): variogram.f90                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int ndim, nstruct;
  okr_struct s[MAXS];
  double mat[MAXS][9];  /* per-structure aniso matrix, variogram.f90:185-192 */
  double cov0;          /* variogram.f90:853-869 */
  /* tables */
  int table_mode, levelB;
  int nz;
  double zb[4], zinv[3];
  int zoff[3], zn[3], ntab;
  double *tabB;         /* Level B composite, 0..ntab */
  double bmat[9];       /* struct_aniso matrix */
  int b_tail;
  double amaj_ref;
  double *tabA[MAXS];   /* Level A per-component tables */
} model_t;

/* variogram.f90:194-210 */
static double aniso_h(const double *m, int ndim, const double *lag) {
  if (ndim == 1) return fabs(lag[0] * m[0]);
  if (ndim == 2) {
    double rx = m[0] * lag[0] + m[3] * lag[1];
    double ry = m[1] * lag[0] + m[4] * lag[1];
    return sqrt(rx * rx + ry * ry);
  }
  double rx = m[0] * lag[0] + m[3] * lag[1] + m[6] * lag[2];
  double ry = m[1] * lag[0] + m[4] * lag[1] + m[7] * lag[2];
  double rz = m[2] * lag[0] + m[5] * lag[1] + m[8] * lag[2];
  return sqrt(rx * rx + ry * ry + rz * rz);
}

/* variogram.f90:327-337 */
static double comp_cov_h(const okr_struct *c, double h) {
  const double eps = DBL_MIN * 1.0e3;
  return (h > eps) ? c->sill * corefunc(c->vtype, h) : c->sill + c->nugget;
}

/* variogram.f90:223-258 zones for h_bounds=[0,.1,.5,3.5], dh=[1e-5,1e-4,1e-3]
 * (kriging.F90:416-417) */
static void setup_zones(model_t *M) {
  const double hb[4] = {0.0, 0.1, 0.5, 3.5};
  const double dh[3] = {1e-5, 1e-4, 1e-3};
  M->nz = 3;
  int off = 0;
  for (int k = 0; k < 3; ++k) {
    double r = (hb[k + 1] - hb[k]) / dh[k];
    int n = (int)lround(r);
    if (n < 1) n = 1;
    M->zn[k] = n;
    M->zoff[k] = off;
    M->zinv[k] = (double)n / (hb[k + 1] - hb[k]);
    off += n;
  }
  for (int k = 0; k < 4; ++k) M->zb[k] = hb[k];
  M->ntab = off;
}

/* variogram.f90:682-710 (and comp_cov_tab_h :346-382) */
static double tab_lookup(const model_t *M, const double *tab, double h, int tail,
                         double (*tailf)(const model_t *, int, double), int ic) {
  if (h >= M->zb[3]) return tail ? tailf(M, ic, h) : 0.0;
  if (h <= 0.0) return tab[0];
  for (int k = 1; k <= M->nz; ++k) {
    if (h < M->zb[k]) {
      double fi = (h - M->zb[k - 1]) * M->zinv[k - 1];
      int ifi = (int)fi;
      int i = M->zoff[k - 1] + (ifi < M->zn[k - 1] - 1 ? ifi : M->zn[k - 1] - 1);
      fi = fi - (double)ifi;
      return tab[i] + fi * (tab[i + 1] - tab[i]);
    }
  }
  return tail ? tailf(M, ic, h) : 0.0;
}

/* struct_cov_h: variogram.f90:712-728 */
static double struct_cov_h(const model_t *M, double h) {
  double res = 0.0;
  int iv = 0;
  while (iv < M->nstruct) {
    double g = comp_cov_h(&M->s[iv], h);
    while (iv + 1 < M->nstruct && M->s[iv + 1].is_product) {
      ++iv;
      g = g * comp_cov_h(&M->s[iv], h);
    }
    res = res + g;
    ++iv;
  }
  return res;
}
static double tail_struct(const model_t *M, int ic, double h) { (void)ic; return struct_cov_h(M, h); }
static double tail_comp(const model_t *M, int ic, double h) { return comp_cov_h(&M->s[ic], h); }

static int model_build(model_t *M, const okr_problem *P, char *err) {
  memset(M, 0, sizeof *M);
  M->ndim = P->ndim;
  M->nstruct = P->nstruct;
  M->table_mode = P->table_mode;
  if (P->nstruct < 1 || P->nstruct > MAXS) { snprintf(err, 512, "bad nstruct"); return 1; }
  for (int i = 0; i < P->nstruct; ++i) {
    M->s[i] = P->structs[i];
    const okr_struct *c = &M->s[i];
    double R[9];
    calc_rotmat(c->azimuth, c->dip, c->plunge, c->a_minor1 / c->a_major, c->a_minor2 / c->a_major, R);
    for (int k = 0; k < 9; ++k) M->mat[i][k] = R[k] / c->a_major;
  }
  /* cov0: variogram.f90:853-869 */
  double tot = 0.0;
  int iv = 0;
  while (iv < M->nstruct) {
    double g = M->s[iv].sill + M->s[iv].nugget;
    while (iv + 1 < M->nstruct && M->s[iv + 1].is_product) {
      ++iv;
      g = g * (M->s[iv].sill + M->s[iv].nugget);
    }
    tot = tot + g;
    ++iv;
  }
  M->cov0 = tot;
  if (!M->table_mode) return 0;
  setup_zones(M);
  /* Level B eligibility: variogram.f90:497-539 */
  const double MAT_TOL = 1.0e-6, ANG_TOL = 1.0e-4;
  int ref = -1, levelB = 1;
  for (int i = 0; i < M->nstruct; ++i) {
    if (M->s[i].vtype == 0) continue;
    if (ref < 0) { ref = i; continue; }
    if (M->s[i].is_product) {
      if (fabs(M->s[i].azimuth - M->s[ref].azimuth) > ANG_TOL ||
          fabs(M->s[i].dip - M->s[ref].dip) > ANG_TOL ||
          fabs(M->s[i].plunge - M->s[ref].plunge) > ANG_TOL) { levelB = 0; break; }
    } else {
      double mx = 0.0;
      for (int k = 0; k < 9; ++k) { double d = fabs(M->mat[i][k] - M->mat[ref][k]); if (d > mx) mx = d; }
      if (mx > MAT_TOL) { levelB = 0; break; }
    }
  }
  if (levelB && ref < 0) { /* all nugget: no table, analytic cov_lag is used (tab not ready) */
    M->levelB = 0;
    M->table_mode = 0;
    return 0;
  }
  M->levelB = levelB;
  if (levelB) {
    /* build_composite_table: variogram.f90:572-680 */
    memcpy(M->bmat, M->mat[ref], sizeof M->bmat);
    M->amaj_ref = M->s[ref].a_major;
    M->tabB = (double *)malloc(sizeof(double) * (M->ntab + 1));
    M->tabB[0] = M->cov0;
    for (int k = 0; k < M->nz; ++k) {
      double hlo = M->zb[k], step = 1.0 / M->zinv[k];
      for (int i = 1; i <= M->zn[k]; ++i) {
        double rdist = hlo + (double)i * step, val = 0.0;
        int j = 0;
        while (j < M->nstruct) {
          if (M->s[j].vtype == 0) { ++j; continue; }
          double rk = rdist * M->amaj_ref / M->s[j].a_major;
          double g = M->s[j].sill * corefunc(M->s[j].vtype, rk);
          while (j + 1 < M->nstruct && M->s[j + 1].is_product) {
            ++j;
            rk = rdist * M->amaj_ref / M->s[j].a_major;
            g = g * M->s[j].sill * corefunc(M->s[j].vtype, rk);
          }
          val = val + g;
          ++j;
        }
        M->tabB[M->zoff[k] + i] = val;
      }
    }
    /* analytic tail flag: variogram.f90:645-672 */
    M->b_tail = 0;
    int j = 0;
    while (j < M->nstruct) {
      if (has_tail(M->s[j].vtype) && M->s[j].vtype != 0) {
        int gt = has_tail(M->s[j].vtype);
        while (j + 1 < M->nstruct && M->s[j + 1].is_product) { ++j; gt = gt && has_tail(M->s[j].vtype); }
        M->b_tail = M->b_tail || gt;
      }
      ++j;
    }
  } else {
    /* Level A: comp_build_table variogram.f90:263-325 */
    for (int c = 0; c < M->nstruct; ++c) {
      double *t = (double *)malloc(sizeof(double) * (M->ntab + 1));
      t[0] = M->s[c].sill + M->s[c].nugget;
      for (int k = 0; k < M->nz; ++k) {
        double hlo = M->zb[k], step = 1.0 / M->zinv[k];
        for (int i = 1; i <= M->zn[k]; ++i) {
          double rdist = hlo + (double)i * step;
          t[M->zoff[k] + i] = M->s[c].sill * corefunc(M->s[c].vtype, rdist);
        }
      }
      M->tabA[c] = t;
    }
  }
  return 0;
}

static void model_free(model_t *M) {
  free(M->tabB);
  for (int i = 0; i < MAXS; ++i) free(M->tabA[i]);
}

/* COV(vgm, lag): cov_dispatch.fh:14-19 -> struct_cov_tab variogram.f90:750-771
 * or struct_cov_lag :731-747 */
static double cov_lag(const model_t *M, const double *lag) {
  if (M->table_mode) {
    if (M->levelB) {
      double h = aniso_h(M->bmat, M->ndim, lag);
      return tab_lookup(M, M->tabB, h, M->b_tail, tail_struct, 0);
    }
    double res = 0.0;
    int iv = 0;
    while (iv < M->nstruct) {
      double h = aniso_h(M->mat[iv], M->ndim, lag);
      double g = tab_lookup(M, M->tabA[iv], h, has_tail(M->s[iv].vtype), tail_comp, iv);
      while (iv + 1 < M->nstruct && M->s[iv + 1].is_product) {
        ++iv;
        h = aniso_h(M->mat[iv], M->ndim, lag);
        g = g * tab_lookup(M, M->tabA[iv], h, has_tail(M->s[iv].vtype), tail_comp, iv);
      }
      res = res + g;
      ++iv;
    }
    return res;
  }
  double res = 0.0;
  int iv = 0;
  while (iv < M->nstruct) {
    double g = comp_cov_h(&M->s[iv], aniso_h(M->mat[iv], M->ndim, lag));
    while (iv + 1 < M->nstruct && M->s[iv + 1].is_product) {
      ++iv;
      g = g * comp_cov_h(&M->s[iv], aniso_h(M->mat[iv], M->ndim, lag));
    }
    res = res + g;
    ++iv;
  }
  return res;
}

/* ------------------------------------------------------------------------ */
/* kd-tree: src/libkriging/kdtree2_maxidx.f90                                */
/* ------------------------------------------------------------------------ */
typedef struct { double lo, hi; } ivl_t;
typedef struct {
  int cut_dim;                /* -1 for a bucket */
  double cut_val, cut_left, cut_right;
  int l, u;                   /* 0-based inclusive range into ind[] */
  int left, right;            /* child node ids or -1 */
  ivl_t box[4];
} knode_t;
typedef struct {
  int dim, n;
  const double *data;  /* (dim, n) */
  int *ind;            /* position -> 0-based point index */
  double *rdata;       /* rearranged (dim, n) */
  knode_t *nodes;
  int nn, cap, root;
} ktree_t;

/* :951-997 (min/max are order independent) */
static ivl_t spread(const ktree_t *t, int c, int l, int u) {
  ivl_t r;
  r.lo = r.hi = t->data[c + t->dim * t->ind[l]];
  for (int i = l + 1; i <= u; ++i) {
    double v = t->data[c + t->dim * t->ind[i]];
    if (v < r.lo) r.lo = v;
    if (v > r.hi) r.hi = v;
  }
  return r;
}

/* :896-914 */
static int select_on_value(ktree_t *t, int c, double alpha, int li, int ui) {
  int lb = li, rb = ui;
  while (lb < rb) {
    if (t->data[c + t->dim * t->ind[lb]] <= alpha) {
      ++lb;
    } else {
      int tmp = t->ind[lb]; t->ind[lb] = t->ind[rb]; t->ind[rb] = tmp;
      --rb;
    }
  }
  return (t->data[c + t->dim * t->ind[lb]] <= alpha) ? lb : lb - 1;
}

static int new_node(ktree_t *t) {
  if (t->nn == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 256;
    t->nodes = (knode_t *)realloc(t->nodes, sizeof(knode_t) * t->cap);
  }
  memset(&t->nodes[t->nn], 0, sizeof(knode_t));
  return t->nn++;
}

/* build_tree_for_range :719-855 (mean split on the widest dimension, bucket
 * when u-l <= 12, boxes of non-cut dimensions inherited from the parent) */
static int build_range(ktree_t *t, int l, int u, int parent, int depth) {
  if (u < l) return -1;
  if (depth > 200) return -2;
  int id = new_node(t);
  knode_t nd;
  memset(&nd, 0, sizeof nd);
  nd.left = nd.right = -1;
  if (u - l <= 12) {
    for (int i = 0; i < t->dim; ++i) nd.box[i] = spread(t, i, l, u);
    nd.cut_dim = -1;
    nd.cut_val = 0.0;
    nd.l = l; nd.u = u;
    t->nodes[id] = nd;
    return id;
  }
  for (int i = 0; i < t->dim; ++i) {
    int recompute = 1;
    if (parent >= 0 && i != t->nodes[parent].cut_dim) recompute = 0;
    nd.box[i] = recompute ? spread(t, i, l, u) : t->nodes[parent].box[i];
  }
  int c = 0;
  double best = nd.box[0].hi - nd.box[0].lo;
  for (int i = 1; i < t->dim; ++i) {
    double w = nd.box[i].hi - nd.box[i].lo;
    if (w > best) { best = w; c = i; }
  }
  double sum = 0.0;
  for (int j = l; j <= u; ++j) sum = sum + t->data[c + t->dim * t->ind[j]];
  double average = sum / (double)(u - l + 1);
  nd.cut_val = average;
  int m = select_on_value(t, c, average, l, u);
  nd.cut_dim = c;
  nd.l = l; nd.u = u;
  t->nodes[id] = nd;
  int left = build_range(t, l, m, id, depth + 1);
  if (left == -2) return -2;
  int right = build_range(t, m + 1, u, id, depth + 1);
  if (right == -2) return -2;
  knode_t *p = &t->nodes[id];
  p->left = left;
  p->right = right;
  if (right < 0) {
    memcpy(p->box, t->nodes[left].box, sizeof p->box);
    p->cut_left = t->nodes[left].box[c].hi;
    p->cut_val = p->cut_left;
  } else if (left < 0) {
    memcpy(p->box, t->nodes[right].box, sizeof p->box);
    p->cut_right = t->nodes[right].box[c].lo;
    p->cut_val = p->cut_right;
  } else {
    p->cut_right = t->nodes[right].box[c].lo;
    p->cut_left = t->nodes[left].box[c].hi;
    p->cut_val = (p->cut_left + p->cut_right) / 2;
    for (int i = 0; i < t->dim; ++i) {
      double uh = t->nodes[left].box[i].hi, ur = t->nodes[right].box[i].hi;
      double ll = t->nodes[left].box[i].lo, lr = t->nodes[right].box[i].lo;
      p->box[i].hi = uh > ur ? uh : ur;
      p->box[i].lo = ll < lr ? ll : lr;
    }
  }
  return id;
}

static int tree_build(ktree_t *t, int dim, int n, const double *data) {
  memset(t, 0, sizeof *t);
  t->dim = dim;
  t->n = n;
  t->data = data;
  t->ind = (int *)malloc(sizeof(int) * n);
  for (int j = 0; j < n; ++j) t->ind[j] = j;
  t->root = build_range(t, 0, n - 1, -1, 0);
  if (t->root == -2) return 1;
  t->rdata = (double *)malloc(sizeof(double) * dim * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < dim; ++k) t->rdata[k + dim * i] = data[k + dim * t->ind[i]];
  return 0;
}

static void tree_free(ktree_t *t) {
  free(t->ind);
  free(t->rdata);
  free(t->nodes);
}

typedef struct { double dis; int idx; } kres_t;
typedef struct {
  const ktree_t *t;
  const double *qv;
  int nn, nfound;
  double ballsize;
  kres_t *heap;  /* 1-based heap positions stored at heap[pos-1] */
  int heap_size;
  int max_idx;   /* only points with 1-based index <= max_idx (kdtree2_n_nearest_maxidx) */
} srec_t;

/* pq_insert :307-346 */
static double pq_insert(srec_t *s, double dis, int idx) {
  kres_t *a = s->heap - 1;
  s->heap_size += 1;
  int i = s->heap_size;
  while (i > 1) {
    int ip = i / 2;
    double pd = a[ip].dis;
    if (dis > pd) {
      a[i] = a[ip];
      i = ip;
    } else {
      break;
    }
  }
  a[i].dis = dis;
  a[i].idx = idx;
  return a[1].dis;
}

/* pq_replace_max :392-446 */
static double pq_replace_max(srec_t *s, double dis, int idx) {
  kres_t *a = s->heap - 1;
  int N = s->heap_size;
  if (N >= 1) {
    int parent = 1, child = 2;
    while (child <= N) {
      double pc = a[child].dis;
      if (child < N) {
        double pc1 = a[child + 1].dis;
        if (pc < pc1) { child = child + 1; pc = pc1; }
      }
      if (dis >= pc) break;
      a[parent] = a[child];
      parent = child;
      child = 2 * parent;
    }
    a[parent].dis = dis;
    a[parent].idx = idx;
    return a[1].dis;
  }
  a[1].dis = dis;
  a[1].idx = idx;
  return dis;
}

/* process_terminal_node :1510-1613 (rearranged data, no correlation window) */
static void process_terminal(srec_t *s, const knode_t *nd) {
  const ktree_t *t = s->t;
  double ballsize = s->ballsize;
  for (int i = nd->l; i <= nd->u; ++i) {
    /* index filter first: process_terminal_node_maxidx :2062-2108 */
    if (t->ind[i] + 1 > s->max_idx) continue;
    double sd = 0.0;
    int skip = 0;
    for (int k = 0; k < t->dim; ++k) {
      double d = t->rdata[k + t->dim * i] - s->qv[k];
      sd = sd + d * d;
      if (sd > ballsize) { skip = 1; break; }
    }
    if (skip) continue;
    int indexofi = t->ind[i] + 1;
    if (s->nfound < s->nn) {
      s->nfound += 1;
      double newpri = pq_insert(s, sd, indexofi);
      if (s->nfound == s->nn) ballsize = newpri;
    } else {
      ballsize = pq_replace_max(s, sd, indexofi);
    }
  }
  s->ballsize = ballsize;
}

static double dis2_from_bnd(double x, double amin, double amax) {
  if (x > amax) { double d = x - amax; return d * d; }
  if (x < amin) { double d = amin - x; return d * d; }
  return 0.0;
}

/* search :1383-1457 */
static void search(srec_t *s, int id) {
  const knode_t *nd = &s->t->nodes[id];
  if (nd->left < 0 || nd->right < 0) {
    process_terminal(s, nd);
    return;
  }
  int c = nd->cut_dim;
  double qval = s->qv[c], dis;
  int closer, farther;
  if (qval < nd->cut_val) {
    closer = nd->left; farther = nd->right;
    double d = nd->cut_right - qval; dis = d * d;
  } else {
    closer = nd->right; farther = nd->left;
    double d = nd->cut_left - qval; dis = d * d;
  }
  if (closer >= 0) search(s, closer);
  if (farther >= 0) {
    double ballsize = s->ballsize;
    if (dis <= ballsize) {
      for (int i = 0; i < s->t->dim; ++i) {
        if (i != c) {
          dis = dis + dis2_from_bnd(s->qv[i], nd->box[i].lo, nd->box[i].hi);
          if (dis > ballsize) return;
        }
      }
      search(s, farther);
    }
  }
}

/* heapsort_struct :1860-1914 (unstable, ascending) */
static void heapsort_struct(kres_t *a1, int n) {
  kres_t *a = a1 - 1;
  if (n <= 1) return;
  int ileft = n / 2 + 1, iright = n;
  kres_t value;
  for (;;) {
    if (ileft > 1) {
      ileft = ileft - 1;
      value = a[ileft];
    } else {
      value = a[iright];
      a[iright] = a[1];
      iright = iright - 1;
      if (iright == 1) { a[1] = value; return; }
    }
    int i = ileft, j = 2 * ileft;
    while (j <= iright) {
      if (j < iright && a[j].dis < a[j + 1].dis) j = j + 1;
      if (value.dis < a[j].dis) {
        a[i] = a[j];
        i = j;
        j = j + j;
      } else {
        j = iright + 1;
      }
    }
    a[i] = value;
  }
}

/* kdtree2_n_nearest :1043-1084 */
static void knn(const ktree_t *t, const double *q, int nn, int sort, kres_t *res) {
  srec_t s;
  s.t = t;
  s.qv = q;
  s.nn = nn;
  s.nfound = 0;
  s.ballsize = DBL_MAX;
  s.heap = res;
  s.heap_size = 0;
  s.max_idx = INT_MAX;
  search(&s, t->root);
  if (sort && nn > 1) heapsort_struct(res, nn);
}

/* kdtree2_n_nearest_maxidx :1928-1964 (the caller guarantees at least nn
 * points below the limit, so the heap fills) */
static void knn_maxidx(const ktree_t *t, const double *q, int nn, int sort, int max_idx, kres_t *res) {
  srec_t s;
  s.t = t;
  s.qv = q;
  s.nn = nn;
  s.nfound = 0;
  s.ballsize = DBL_MAX;
  s.heap = res;
  s.heap_size = 0;
  s.max_idx = max_idx;
  search(&s, t->root);
  if (sort && nn > 1) heapsort_struct(res, nn);
}

int okr_knn_maxidx(int ndim, int n, const double *coord, int nq, const double *q, const int *k, int sort,
                   const int *max_idx, int kmax, int *idx, double *dis) {
  ktree_t t;
  if (tree_build(&t, ndim, n, coord)) { tree_free(&t); return 1; }
  kres_t *r = (kres_t *)malloc(sizeof(kres_t) * (kmax > 0 ? kmax : 1));
  for (int j = 0; j < nq; ++j) {
    knn_maxidx(&t, q + (size_t)ndim * j, k[j], sort, max_idx[j], r);
    for (int i = 0; i < k[j]; ++i) { idx[(size_t)kmax * j + i] = r[i].idx; dis[(size_t)kmax * j + i] = r[i].dis; }
  }
  free(r);
  tree_free(&t);
  return 0;
}

int okr_knn(int ndim, int n, const double *coord, const double *q, int k, int sort, int *idx,
            double *dis) {
  ktree_t t;
  if (tree_build(&t, ndim, n, coord)) { tree_free(&t); return 1; }
  kres_t *r = (kres_t *)malloc(sizeof(kres_t) * k);
  knn(&t, q, k, sort, r);
  for (int i = 0; i < k; ++i) { idx[i] = r[i].idx; dis[i] = r[i].dis; }
  free(r);
  tree_free(&t);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* dense helpers                                                             */
/* ------------------------------------------------------------------------ */
/* upper Cholesky K = U^T U in place (column-major, ld = n); info like SPOTRF */
static int chol_upper(double *A, int n, int ld) {
  for (int j = 0; j < n; ++j) {
    double s = A[j + ld * j];
    for (int k = 0; k < j; ++k) s -= A[k + ld * j] * A[k + ld * j];
    if (!(s > 0.0)) return j + 1;
    double d = sqrt(s);
    A[j + ld * j] = d;
    for (int i = j + 1; i < n; ++i) {
      double v = A[j + ld * i];
      for (int k = 0; k < j; ++k) v -= A[k + ld * j] * A[k + ld * i];
      A[j + ld * i] = v / d;
    }
  }
  return 0;
}
/* solve U^T U x = b for one rhs (SPOTRS 'U') */
static void chol_solve(const double *U, int n, int ld, double *b) {
  for (int i = 0; i < n; ++i) {
    double v = b[i];
    for (int k = 0; k < i; ++k) v -= U[k + ld * i] * b[k];
    b[i] = v / U[i + ld * i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double v = b[i];
    for (int k = i + 1; k < n; ++k) v -= U[i + ld * k] * b[k];
    b[i] = v / U[i + ld * i];
  }
}
/* Bunch-Kaufman factorisation and solve of the full augmented system: the
 * reference's SSYTRF/SSYTRS fallback (solver.f90:337-406), restated from the
 * vendored LAPACK, src/libkriging/lapack.f: SSYTF2 upper (:6213-6445; SSYTRF
 * runs it for every system of up to 64 rows, :6811-6900), ISAMAX (:669-720),
 * SSYR, SGER, SGEMV and SSCAL in their reference loop orders, and SSYTRS
 * upper (:7478-7600).  Column-major, leading dimension ld, 0-based indices;
 * ipiv holds LAPACK's 1-based codes (k+1 for a 1x1 pivot, -(kp+1) on both
 * columns of a 2x2 pivot).  The return value is INFO: 0, or the 1-based
 * column whose pivot block is exactly zero. */
static int iamax(int n, const double *x, int inc) {
  int im = 0;
  double mx = fabs(x[0]);
  for (int i = 1; i < n; ++i) {
    double v = fabs(x[(size_t)i * inc]);
    if (v > mx) { mx = v; im = i; }
  }
  return im;
}
static int bk_factor(double *A, int n, int ld, int *ipiv) {
  const double alpha = (1.0 + sqrt(17.0)) / 8.0;
  int info = 0;
  int k = n - 1;
  while (k >= 0) {
    int kstep = 1, kp, imax = 0;
    double absakk = fabs(A[k + ld * k]), colmax = 0.0;
    if (k > 0) {
      imax = iamax(k, A + (size_t)ld * k, 1);
      colmax = fabs(A[imax + ld * k]);
    }
    double mx = absakk > colmax ? absakk : colmax;
    if (mx == 0.0 || isnan(absakk)) {
      if (info == 0) info = k + 1;
      kp = k;
    } else {
      if (absakk >= alpha * colmax) {
        kp = k;
      } else {
        int jmax = imax + 1 + iamax(k - imax, A + imax + (size_t)ld * (imax + 1), ld);
        double rowmax = fabs(A[imax + ld * jmax]);
        if (imax > 0) {
          jmax = iamax(imax, A + (size_t)ld * imax, 1);
          double v = fabs(A[jmax + ld * imax]);
          rowmax = rowmax > v ? rowmax : v;
        }
        if (absakk >= alpha * colmax * (colmax / rowmax)) kp = k;
        else if (fabs(A[imax + ld * imax]) >= alpha * rowmax) kp = imax;
        else { kp = imax; kstep = 2; }
      }
      int kk = k - kstep + 1;
      if (kp != kk) {
        for (int i = 0; i < kp; ++i) { double t = A[i + ld * kk]; A[i + ld * kk] = A[i + ld * kp]; A[i + ld * kp] = t; }
        for (int i = kp + 1; i < kk; ++i) { double t = A[i + ld * kk]; A[i + ld * kk] = A[kp + ld * i]; A[kp + ld * i] = t; }
        double t = A[kk + ld * kk]; A[kk + ld * kk] = A[kp + ld * kp]; A[kp + ld * kp] = t;
        if (kstep == 2) { t = A[(k - 1) + ld * k]; A[(k - 1) + ld * k] = A[kp + ld * k]; A[kp + ld * k] = t; }
      }
      if (kstep == 1) {
        /* SSYR upper with alpha = -1/d, then SSCAL of the column */
        double r1 = 1.0 / A[k + ld * k];
        for (int j = 0; j < k; ++j) {
          double xj = A[j + ld * k];
          if (xj != 0.0) {
            double tmp = -r1 * xj;
            for (int i = 0; i <= j; ++i) A[i + ld * j] = A[i + ld * j] + A[i + ld * k] * tmp;
          }
        }
        for (int i = 0; i < k; ++i) A[i + ld * k] = r1 * A[i + ld * k];
      } else if (k > 1) {
        double d12 = A[(k - 1) + ld * k];
        double d22 = A[(k - 1) + ld * (k - 1)] / d12;
        double d11 = A[k + ld * k] / d12;
        double t = 1.0 / (d11 * d22 - 1.0);
        d12 = t / d12;
        for (int j = k - 2; j >= 0; --j) {
          double wkm1 = d12 * (d11 * A[j + ld * (k - 1)] - A[j + ld * k]);
          double wk = d12 * (d22 * A[j + ld * k] - A[j + ld * (k - 1)]);
          for (int i = j; i >= 0; --i) A[i + ld * j] = A[i + ld * j] - A[i + ld * k] * wk - A[i + ld * (k - 1)] * wkm1;
          A[j + ld * k] = wk;
          A[j + ld * (k - 1)] = wkm1;
        }
      }
    }
    if (kstep == 1) ipiv[k] = kp + 1;
    else { ipiv[k] = -(kp + 1); ipiv[k - 1] = -(kp + 1); }
    k -= kstep;
  }
  return info;
}
/* SSYTRS upper for nrhs right-hand sides B (column-major, leading dimension ldb) */
static void bk_solve(const double *A, int n, int ld, const int *ipiv, double *B, int nrhs, int ldb) {
#define SWAPROW(r1, r2) for (int j_ = 0; j_ < nrhs; ++j_) { double t_ = B[(r1) + ldb * j_]; B[(r1) + ldb * j_] = B[(r2) + ldb * j_]; B[(r2) + ldb * j_] = t_; }
  int k = n - 1;
  while (k >= 0) {
    if (ipiv[k] > 0) {
      int kp = ipiv[k] - 1;
      if (kp != k) SWAPROW(k, kp);
      for (int j = 0; j < nrhs; ++j) {  /* SGER */
        double y = B[k + ldb * j];
        if (y != 0.0) {
          double tmp = -y;
          for (int i = 0; i < k; ++i) B[i + ldb * j] = B[i + ldb * j] + A[i + ld * k] * tmp;
        }
      }
      double r = 1.0 / A[k + ld * k];  /* SSCAL */
      for (int j = 0; j < nrhs; ++j) B[k + ldb * j] = r * B[k + ldb * j];
      k -= 1;
    } else {
      int kp = -ipiv[k] - 1;
      if (kp != k - 1) SWAPROW(k - 1, kp);
      for (int c = 0; c < 2; ++c) {  /* SGER with columns k, then k-1 */
        int col = k - c;
        for (int j = 0; j < nrhs; ++j) {
          double y = B[col + ldb * j];
          if (y != 0.0) {
            double tmp = -y;
            for (int i = 0; i < k - 1; ++i) B[i + ldb * j] = B[i + ldb * j] + A[i + ld * col] * tmp;
          }
        }
      }
      double akm1k = A[(k - 1) + ld * k];
      double akm1 = A[(k - 1) + ld * (k - 1)] / akm1k;
      double ak = A[k + ld * k] / akm1k;
      double denom = akm1 * ak - 1.0;
      for (int j = 0; j < nrhs; ++j) {
        double bkm1 = B[(k - 1) + ldb * j] / akm1k;
        double bk = B[k + ldb * j] / akm1k;
        B[(k - 1) + ldb * j] = (ak * bkm1 - bk) / denom;
        B[k + ldb * j] = (akm1 * bk - bkm1) / denom;
      }
      k -= 2;
    }
  }
  k = 0;
  while (k < n) {
    int two = ipiv[k] < 0;
    for (int c = 0; c <= two; ++c) {  /* SGEMV 'T', alpha -1, beta 1 */
      int col = k + c;
      for (int j = 0; j < nrhs; ++j) {
        double tmp = 0.0;
        for (int i = 0; i < k; ++i) tmp = tmp + B[i + ldb * j] * A[i + ld * col];
        if (k > 0) B[col + ldb * j] = B[col + ldb * j] + (-1.0) * tmp;
      }
    }
    int kp = (two ? -ipiv[k] : ipiv[k]) - 1;
    if (kp != k) SWAPROW(k, kp);
    k += two ? 2 : 1;
  }
#undef SWAPROW
}
/* the reference's fallback for one augmented system of m rows (n data rows):
 * SSYTRF + SSYTRS, and when SSYTRF reports an exactly singular block one
 * retry with 1e-6 * mean(data diagonal) added to the data diagonal
 * (kriging_base.F90:3299-3345).  K (ld = m) is modified by the jitter as the
 * reference's matA is; B holds nrhs right-hand sides (ldb = m) on entry and
 * the solutions on success.  Returns INFO. */
static int bk_fallback(double *K, int n, int m, double *B, int nrhs) {
  double *A = (double *)malloc(sizeof(double) * m * m);
  double *B0 = (double *)malloc(sizeof(double) * m * nrhs);
  int *ipiv = (int *)malloc(sizeof(int) * m);
  memcpy(B0, B, sizeof(double) * m * nrhs);
  memcpy(A, K, sizeof(double) * m * m);
  int info = bk_factor(A, m, m, ipiv);
  if (info != 0 && n > 0) {
    double ds = 0.0;
    for (int i = 0; i < n; ++i) ds = ds + K[i + m * i];
    if (ds > 0.0) {
      ds = 1.0e-6 * (ds / (double)n);
      for (int i = 0; i < n; ++i) K[i + m * i] = K[i + m * i] + ds;
      memcpy(A, K, sizeof(double) * m * m);
      info = bk_factor(A, m, m, ipiv);
    }
  }
  if (info == 0) bk_solve(A, m, m, ipiv, B, nrhs, m);
  free(A);
  free(B0);
  free(ipiv);
  return info;
}

/* exported for the numpy cokriging restatement: the same fallback on one
 * column-major m x m system K (modified by the jitter) with nrhs columns B */
int okr_bk_fallback(double *K, int n, int m, double *B, int nrhs) { return bk_fallback(K, n, m, B, nrhs); }

/* insertion sort: kriging_base.F90:3510-3524 */
static void isort(int *a, int n) {
  for (int i = 1; i < n; ++i) {
    int tmp = a[i], j = i - 1;
    while (j >= 0 && a[j] > tmp) { a[j + 1] = a[j]; --j; }
    a[j + 1] = tmp;
  }
}

/* kriging_normalize_nmax: kriging_base.F90:1966-1971 */
static int normalize_nmax(int nmax, int ntot) {
  int v = nmax < ntot ? nmax : ntot;
  if (v <= 0) v = ntot;
  return v;
}

static int eff_nmax(const okr_problem *P) {
  int nm = P->nmax;
  /* set_grid_cv_base: kriging_base.F90:1167 (nmax+1 on the raw value; the
   * INT_MAX default wraps negative and then normalises to all obs) */
  if (P->cross_validation && nm > 0) nm = (int)((unsigned)nm + 1u);
  return normalize_nmax(nm, P->nobs);
}

int okr_mmax(const okr_problem *P) {
  int nm = eff_nmax(P);
  int need = P->nobs > nm;
  if (P->sector && need) return (1 << P->ndim) * nm;
  return nm;
}

/* ------------------------------------------------------------------------ */
/* one block                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  const okr_problem *P;
  const model_t *M;
  const ktree_t *tree;  /* NULL when every obs is used */
  int nmax, need_search, ncand, mmax;
  double maxdist2;
  double srot[9];
  int have_srot;
  /* space-time */
  int st, nl;           /* nl: coordinate rows (ndim, or 4 for ST) */
  const model_t *Mt;    /* temporal structures */
  double k_ps, cov0;    /* cov0 of the (ST) model */
  double time_at;
  /* sum-metric: joint sills and the temporal transform */
  int st_sum, t_vtype;
  double t_nug, t_sill, t_at;
  const double *sill_st;
} ctx_t;

/* f_time_vgm_st, variogram_st.f90:126-141: f(0) = 0 (|dt| <= epsilon),
 * otherwise max(nugget + sill (1 - corefunc(|dt| / at)), 0) */
static double f_time(const ctx_t *C, double dt) {
  if (fabs(dt) <= DBL_EPSILON) return 0.0;
  double adt = fabs(dt) / C->t_at;
  double dw = C->t_nug + C->t_sill * (1.0 - corefunc(C->t_vtype, adt));
  return dw > 0.0 ? dw : 0.0;
}

/* COV / COV_ST: cov_lag or, for space-time, the product-sum
 * k_ps Cs(lag_s) Ct(|dt|) + Cs(lag_s) + Ct(|dt|) (variogram_st.f90:222-282;
 * the temporal lag is the 1-D vector [|dt|, 0, 0]) or the sum-metric
 * Cs + Ct + sum_k sill_st[k] shape_k(h_st), h_st = sqrt(h_k^2 + f(dt)^2) with
 * h_k structure k's anisotropic lag (variogram_st.f90:178-194 analytic,
 * 241-261 tables: the shape from the structure's own table divided by its
 * sill when the spatial model has per-structure tables and the sill is
 * positive, analytic otherwise) */
static double ctx_cov(const ctx_t *C, const double *lag) {
  if (!C->st) return cov_lag(C->M, lag);
  double ls[3] = {lag[0], lag[1], lag[2]};
  double lt[3] = {fabs(lag[3]), 0.0, 0.0};
  double cs = cov_lag(C->M, ls), ct = cov_lag(C->Mt, lt);
  if (!C->st_sum) return C->k_ps * cs * ct + cs + ct;
  const model_t *M = C->M;
  double res = cs + ct;
  double dw = f_time(C, lag[3]);
  for (int k = 0; k < M->nstruct; ++k) {
    double hs = aniso_h(M->mat[k], 3, ls);
    double hst = sqrt(hs * hs + dw * dw);
    if (M->table_mode && !M->levelB && M->tabA[k] && M->s[k].sill > 0.0)
      res = res + C->sill_st[k] * tab_lookup(M, M->tabA[k], hst, has_tail(M->s[k].vtype), tail_comp, k) / M->s[k].sill;
    else
      res = res + C->sill_st[k] * corefunc(M->s[k].vtype, hst);
  }
  return res;
}

/* compute_cov0_vgm_st, variogram_st.f90:344-375 */
static double st_cov0(const ctx_t *C) {
  if (!C->st_sum) return C->k_ps * C->M->cov0 * C->Mt->cov0 + C->M->cov0 + C->Mt->cov0;
  double c0 = C->M->cov0 + C->Mt->cov0;
  for (int k = 0; k < C->M->nstruct; ++k)
    if (C->M->s[k].vtype != 0) c0 = c0 + C->sill_st[k];
  return c0;
}

static void ctx_set_st(ctx_t *C, const okr_problem *P) {
  C->st_sum = P->st_sum;
  C->t_vtype = P->t_vtype;
  C->t_nug = P->t_nugget;
  C->t_sill = P->t_sill;
  C->t_at = P->t_at;
  C->sill_st = P->sill_st;
}

static void rot_pt(const double *R, int ndim, const double *x, double *y) {
  if (ndim == 1) {
    y[0] = R[0] * x[0];
  } else if (ndim == 2) {
    y[0] = R[0] * x[0] + R[3] * x[1];
    y[1] = R[1] * x[0] + R[4] * x[1];
  } else {
    y[0] = R[0] * x[0] + R[3] * x[1] + R[6] * x[2];
    y[1] = R[1] * x[0] + R[4] * x[1] + R[7] * x[2];
    y[2] = R[2] * x[0] + R[5] * x[1] + R[8] * x[2];
  }
}

/* search_neighbors: kriging.F90:490-666 (+ filter_by_maxlag) */
static int block_search(const ctx_t *C, int ib, const double *xloc, int *inear, double *dist,
                        kres_t *cand) {
  const okr_problem *P = C->P;
  int nd = P->ndim, nl = C->nl, nnear = 0;
  double newloc[4];
  if (C->have_srot) rot_pt(C->srot, nd, xloc, newloc);
  else for (int k = 0; k < nd; ++k) newloc[k] = xloc[k];
  /* ST: search coordinate t * time_at (kriging_st.F90:90-95, 490-497) */
  if (C->st) newloc[3] = xloc[3] * C->time_at;
  if (P->sector && C->need_search) {
    knn(C->tree, newloc, C->ncand, 1, cand);
    int cv_skip = P->cross_validation ? ib + 1 : 0;
    int count[8] = {0};
    /* apply_sector_filter kriging.F90:1093-1154 */
    for (int i = 0; i < C->ncand; ++i) {
      int idx = cand[i].idx;
      if (idx == cv_skip) continue;
      if (cand[i].dis > C->maxdist2) continue;
      const double *o = P->obs_coord + nl * (idx - 1);
      double pr[3];
      if (C->have_srot) rot_pt(C->srot, nd, o, pr);
      else for (int k = 0; k < nd; ++k) pr[k] = o[k];
      double dx = pr[0] - newloc[0], dy = pr[1] - newloc[1], dz = 0.0;
      if (nd == 3) dz = pr[2] - newloc[2];
      int ix = dx < 0.0, iy = dy < 0.0, sec = ix + 2 * iy;
      if (nd == 3) sec += 4 * (dz < 0.0);
      if (count[sec] < C->nmax) {
        count[sec] += 1;
        inear[nnear] = idx;
        dist[nnear] = cand[i].dis;
        ++nnear;
      }
    }
  } else {
    if (C->nmax < P->nobs) {
      knn(C->tree, newloc, C->nmax, 0, cand);
      nnear = C->nmax;
      for (int i = 0; i < nnear; ++i) { inear[i] = cand[i].idx; dist[i] = cand[i].dis; }
    } else {
      /* rotated_dists rotation.f90:193-218 */
      nnear = P->nobs;
      for (int i = 0; i < nnear; ++i) {
        const double *o = P->obs_coord + nl * i;
        double df[3], r[3];
        for (int k = 0; k < nd; ++k) df[k] = o[k] - xloc[k];
        if (C->have_srot) rot_pt(C->srot, nd, df, r);
        else for (int k = 0; k < nd; ++k) r[k] = df[k];
        double s = 0.0;
        for (int k = 0; k < nd; ++k) s = s + r[k] * r[k];
        if (C->st) {
          /* kriging_st.F90:609-612: + (t_obs * at - block_ht)^2 */
          double dt = o[3] * C->time_at - newloc[3];
          s = s + dt * dt;
        }
        inear[i] = i + 1;
        dist[i] = s;
      }
    }
    if (P->cross_validation) {
      for (int i = 0; i < nnear; ++i) {
        if (inear[i] == ib + 1) {
          for (int j = i; j < nnear - 1; ++j) { inear[j] = inear[j + 1]; dist[j] = dist[j + 1]; }
          --nnear;
          break;
        }
      }
    }
  }
  /* filter_by_maxlag kriging_base.F90:2491-2505 */
  int kk = 0;
  for (int i = 0; i < nnear; ++i)
    if (dist[i] <= C->maxdist2) { inear[kk] = inear[i]; dist[kk] = dist[i]; ++kk; }
  return kk;
}

static int run_block(const ctx_t *C, int ib, okr_result *R, kres_t *cand, double *work,
                     int *fail) {
  const okr_problem *P = C->P;
  int nd = C->nl;  /* coordinate rows */
  const double *xloc = P->cross_validation ? P->obs_coord + nd * ib : P->block_coord + nd * ib;
  int *inear = R->inear + (size_t)C->mmax * ib;
  double *dist = work;
  int n = block_search(C, ib, xloc, inear, dist, cand);
  R->nnear[ib] = n;
  if (n == 0) {
    /* kriging_base.F90:3198-3207 + :3421-3425 */
    if (!P->neglect_error) return 2;
    R->estimate[ib] = NAN;
    R->variance[ib] = NAN;
    return 0;
  }
  isort(inear, n);
  if (n == 1) {
    /* solve_base (kriging_base.F90:2127) skips the solve when npp == 1: the
     * variance keeps its NaN fill, the estimate would reuse the weights the
     * same OpenMP thread computed for its previous block.  Both are reported
     * as NaN here (the estimate deliberately: the reference value depends on
     * thread scheduling). */
    R->estimate[ib] = NAN;
    R->variance[ib] = NAN;
    return 0;
  }
  /* the ST assembly does not range-scale lags (kriging_st.F90:805, 921) */
  double rs = (P->rangescale && !C->st) ? P->rangescale[ib] : 1.0;
  double ln = P->localnugget ? P->localnugget[ib] : 0.0;
  int p = P->unbias ? 1 : 0, m = n + p;
  double *K = work + C->mmax;            /* m x m */
  double *b = K + (size_t)m * m;          /* m */
  double *x = b + m;                      /* m */
  double *U = x + m;                      /* n x n */
  double *Y = U + (size_t)n * n;          /* n */
  double lag[4] = {0, 0, 0, 0};
  /* assemble_lhs kriging.F90:714-849 */
  for (int i = 0; i < n; ++i) {
    const double *xi = P->obs_coord + nd * (inear[i] - 1);
    double ov = P->obs_var ? P->obs_var[inear[i] - 1] : 0.0;
    K[i + m * i] = C->cov0 + ov + ln;
    for (int j = i + 1; j < n; ++j) {
      const double *xj = P->obs_coord + nd * (inear[j] - 1);
      for (int k = 0; k < nd; ++k) lag[k] = (xi[k] - xj[k]) / rs;
      double c = ctx_cov(C, lag);
      K[j + m * i] = c;
      K[i + m * j] = c;
    }
    if (p) { K[n + m * i] = 1.0; K[i + m * n] = 1.0; }
  }
  if (p) K[n + m * n] = 0.0;
  /* assemble_rhs kriging.F90:864-939 */
  int ns = P->nsub > 0 ? P->nsub : 1;
  for (int i = 0; i < n; ++i) {
    const double *xi = P->obs_coord + nd * (inear[i] - 1);
    double tmp = 0.0;
    for (int s = 0; s < ns; ++s) {
      const double *g = (ns == 1 && !P->sub_coord) ? xloc : P->sub_coord + (size_t)nd * ((size_t)ib * ns + s);
      double w = (ns == 1 && !P->sub_weight) ? 1.0 : P->sub_weight[(size_t)ib * ns + s];
      for (int k = 0; k < nd; ++k) lag[k] = (xi[k] - g[k]) / rs;
      tmp = tmp + ctx_cov(C, lag) * w;
    }
    b[i] = tmp;
  }
  if (p) b[n] = 1.0;
  /* kriging_setup / kriging_solve_prepared: solver.f90:224-300 */
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) U[i + n * j] = K[i + m * j];
  int info = chol_upper(U, n, n);
  if (info == 0) {
    for (int i = 0; i < n; ++i) x[i] = b[i];
    chol_solve(U, n, n, x);
    if (p) {
      for (int i = 0; i < n; ++i) Y[i] = 1.0;
      chol_solve(U, n, n, Y);
      double S = 0.0, ytb = 0.0;
      for (int i = 0; i < n; ++i) { S += Y[i]; ytb += Y[i] * b[i]; }
      if (!(S > 0.0)) info = 1;
      else {
        double lam = (ytb - b[n]) / S;
        for (int i = 0; i < n; ++i) x[i] = x[i] - Y[i] * lam;
        x[n] = lam;
      }
    }
    if (info == 0) R->stats[1] += 1;
  }
  if (info != 0) {
    /* SSYTRF fallback with a single jittered retry: kriging_base.F90:3299-3366 */
    memcpy(x, b, sizeof(double) * m);
    info = bk_fallback(K, n, m, x, 1);
    if (info == 0) R->stats[3] += 1;
  }
  if (info != 0) {
    if (!P->neglect_error) return 3;
    /* x = NaN (kriging_base.F90:3369-3383): the estimate is NaN; the variance
     * max(NaN, 0) of calc_variance (kriging.F90:1019-1025) is 0 */
    *fail += 1;
    R->estimate[ib] = NAN;
    R->variance[ib] = 0.0;
    if (R->weights)
      for (int i = 0; i < n; ++i) R->weights[(size_t)C->mmax * ib + i] = NAN;
    return 0;
  }
  /* calc_variance kriging.F90:968-1028 */
  double base = C->cov0;
  if (ns > 1) {
    base = 0.0;
    for (int i = 0; i < ns; ++i) {
      const double *gi = P->sub_coord + (size_t)nd * ((size_t)ib * ns + i);
      double wi = P->sub_weight[(size_t)ib * ns + i];
      base = base + C->cov0 * wi * wi;
      for (int j = i + 1; j < ns; ++j) {
        const double *gj = P->sub_coord + (size_t)nd * ((size_t)ib * ns + j);
        double wj = P->sub_weight[(size_t)ib * ns + j];
        for (int k = 0; k < nd; ++k) lag[k] = gi[k] - gj[k];
        base = base + ctx_cov(C, lag) * wi * wj * 2.0;
      }
    }
  }
  double dot = 0.0;
  for (int i = 0; i < m; ++i) dot = dot + x[i] * b[i];
  double var = base - dot;
  R->variance[ib] = var > 0.0 ? var : 0.0;
  /* estimate_block_base kriging_base.F90:3407-3503 */
  double avg = 0.0, tw = 0.0;
  for (int i = 0; i < n; ++i) {
    tw = tw + x[i];
    avg = avg + P->obs_value[inear[i] - 1] * x[i];
  }
  if (P->unbias == 0 && P->sk_mean != 0.0) avg = avg + (1.0 - tw) * P->sk_mean;
  if (avg < P->bound_lo) avg = P->bound_lo;
  if (avg > P->bound_hi) avg = P->bound_hi;
  R->estimate[ib] = avg;
  if (R->weights)
    for (int i = 0; i < n; ++i) R->weights[(size_t)C->mmax * ib + i] = x[i];
  return 0;
}

int okr_cov(const okr_problem *P, int n, const double *lags, double *out) {
  model_t M, Mt;
  char err[512];
  if (model_build(&M, P, err)) return 1;
  ctx_t C;
  memset(&C, 0, sizeof C);
  C.M = &M;
  int nl = P->ndim;
  if (P->st) {
    okr_problem Pt = *P;
    Pt.ndim = 1;
    Pt.nstruct = P->nstruct_t;
    Pt.structs = P->structs_t;
    if (model_build(&Mt, &Pt, err)) {
      model_free(&M);
      return 1;
    }
    C.st = 1;
    C.Mt = &Mt;
    C.k_ps = P->k_ps;
    ctx_set_st(&C, P);
    nl = 4;
  }
  for (int i = 0; i < n; ++i) {
    double lag[4] = {0, 0, 0, 0};
    for (int k = 0; k < nl; ++k) lag[k] = lags[k + nl * i];
    out[i] = ctx_cov(&C, lag);
  }
  model_free(&M);
  if (P->st) model_free(&Mt);
  return 0;
}

int okr_solve(const okr_problem *P, okr_result *R) {
  R->errmsg[0] = 0;
  memset(R->stats, 0, sizeof R->stats);
  if (P->ndim < 1 || P->ndim > 3 || P->nobs < 1 || (P->ndim == 1 && P->sector) || (P->st && P->ndim != 3)) {
    snprintf(R->errmsg, sizeof R->errmsg, "bad ndim/nobs (sector search needs ndim >= 2, ST ndim 3)");
    return 1;
  }
  model_t M, Mt;
  if (model_build(&M, P, R->errmsg)) return 1;
  ctx_t C;
  memset(&C, 0, sizeof C);
  memset(&Mt, 0, sizeof Mt);
  C.P = P;
  C.M = &M;
  C.nl = P->ndim;
  C.cov0 = M.cov0;
  if (P->st) {
    /* temporal model: 1-D, isotropic range at_k (variogram_st.f90:302-317);
     * cov0 = k Cs0 Ct0 + Cs0 + Ct0 (variogram_st.f90:344-375) */
    okr_problem Pt = *P;
    Pt.ndim = 1;
    Pt.nstruct = P->nstruct_t;
    Pt.structs = P->structs_t;
    if (model_build(&Mt, &Pt, R->errmsg)) {
      model_free(&M);
      return 1;
    }
    C.st = 1;
    C.nl = 4;
    C.Mt = &Mt;
    C.k_ps = P->k_ps;
    C.time_at = P->time_at;
    ctx_set_st(&C, P);
    C.cov0 = st_cov0(&C);
  }
  C.nmax = eff_nmax(P);
  C.need_search = P->nobs > C.nmax;
  C.maxdist2 = P->maxdist * P->maxdist;
  C.mmax = okr_mmax(P);
  if (R->mmax != C.mmax) {
    snprintf(R->errmsg, sizeof R->errmsg, "mmax mismatch: expected %d", C.mmax);
    model_free(&M);
    return 1;
  }
  C.ncand = C.nmax;
  if (P->sector && C.need_search) {
    long nc = (long)(1 << P->ndim) * C.nmax * 4;
    C.ncand = (int)(nc < P->nobs ? nc : P->nobs);
  }
  /* anisotropic search: kriging.F90:356-357, 371-381 */
  C.have_srot = P->aniso_search && (fabs(P->s_anis1 - 1.0) > EPSLON || fabs(P->s_anis2 - 1.0) > EPSLON);
  double *rc = NULL;
  ktree_t tree;
  memset(&tree, 0, sizeof tree);
  if (C.have_srot) calc_rotmat(P->s_azimuth, P->s_dip, P->s_plunge, P->s_anis1, P->s_anis2, C.srot);
  if (C.need_search) {
    const double *src = P->obs_coord;
    if (C.have_srot || C.st) {
      /* search coordinates: rotated spatial part, ST time * time_at
       * (kriging_st.F90:387-395) */
      rc = (double *)malloc(sizeof(double) * C.nl * P->nobs);
      for (int i = 0; i < P->nobs; ++i) {
        const double *o = P->obs_coord + C.nl * i;
        double *r = rc + C.nl * i;
        if (C.have_srot) rot_pt(C.srot, P->ndim, o, r);
        else for (int k = 0; k < P->ndim; ++k) r[k] = o[k];
        if (C.st) r[3] = o[3] * C.time_at;
      }
      src = rc;
    }
    if (tree_build(&tree, C.nl, P->nobs, src)) {
      snprintf(R->errmsg, sizeof R->errmsg, "kd-tree build did not terminate");
      tree_free(&tree);
      free(rc);
      model_free(&M);
      if (C.st) model_free(&Mt);
      return 1;
    }
    C.tree = &tree;
  }
  int nb = P->cross_validation ? P->nobs : P->nblock;
  int err = 0, nfail = 0, nchol = 0, nsy = 0;
#pragma omp parallel num_threads(P->nthread > 0 ? P->nthread : 1) reduction(+ : nfail, nchol, nsy)
  {
    int mm = C.mmax + 2;
    kres_t *cand = (kres_t *)malloc(sizeof(kres_t) * (C.ncand > C.mmax ? C.ncand : C.mmax) + 16);
    double *work = (double *)malloc(sizeof(double) * ((size_t)C.mmax + 4 * (size_t)mm * mm + 8 * mm));
    okr_result Rl = *R;
    memset(Rl.stats, 0, sizeof Rl.stats);
#pragma omp for schedule(dynamic, 64)
    for (int ib = 0; ib < nb; ++ib) {
      if (err) continue;
      int rc2 = run_block(&C, ib, &Rl, cand, work, &nfail);
      if (rc2) {
#pragma omp critical
        {
          if (!err) {
            err = rc2;
            snprintf(R->errmsg, sizeof R->errmsg,
                     rc2 == 2 ? "not enough neighbors for kriging at block %d" : "Singular matrix at block %d",
                     ib + 1);
          }
        }
      }
    }
    nchol += Rl.stats[1];
    nsy += Rl.stats[3];
    free(cand);
    free(work);
  }
  R->stats[0] = nfail;
  R->stats[1] = nchol;
  R->stats[3] = nsy;
  tree_free(&tree);
  free(rc);
  model_free(&M);
  if (C.st) model_free(&Mt);
  return err ? 1 : 0;
}
