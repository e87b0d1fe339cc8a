/*
 * CPU oracle for the se_e2_a descriptor-to-force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference's CPU semantics, written to check the HIP kernels against.  Only
 * tests/, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product library never links or calls it.
 *
 * Every function cites the reference function it restates (paths relative to
 * the reference repository root).  Floating-point evaluation order follows the
 * reference statement by statement and the file is compiled with
 * -ffp-contract=off, so on x86-64 (no FMA in the baseline ISA) the results are
 * bit-identical to the compiled reference functions.
 *
 * Pinning (see tests/test_oracle_golden.py and tests/test_oracle_ref.py):
 *   normalize/copy_coord/format_nlist : test_fmt_nlist.cc expected ext-index
 *       lists; brute-force candidate search; the compiled reference library.
 *   env_mat_a                         : test_env_mat_a.cc expected_env, central
 *       finite differences of em w.r.t. coordinates, the compiled reference.
 *   tabulate (fwd/grad/grad_grad)     : test_tabulate_se_a.cc and
 *       test_tabulate_fusion_se_a.py arrays, finite differences, reference.
 *   prod_force_a / prod_virial_a      : test_prod_force_a.cc /
 *       test_prod_virial_a.cc expected arrays, sum-of-forces identity, ref.
 *   prod_force_grad_a / prod_virial_grad_a : test_prod_force_grad_a.cc /
 *       test_prod_virial_grad_a.cc expected arrays, adjoint identity with
 *       prod_force_a / prod_virial_a, ref.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Region helpers: source/lib/src/region.cc:35-122 (compute_volume,
 * compute_rec_boxt, tensor_dot_vec, tensor_t_dot_vec) and
 * source/lib/include/SimulationRegion_Impl.h:344-470 (phys2Inter, inter2Phys,
 * toFaceDistance, computeVolume, computeRecBox).  Both classes use the same
 * cofactor formulas; rows of boxt are the cell vectors. */

static double box_volume(const double *b) {
  double v = b[0] * (b[4] * b[8] - b[7] * b[5]) -
             b[1] * (b[3] * b[8] - b[6] * b[5]) +
             b[2] * (b[3] * b[7] - b[6] * b[4]);
  return fabs(v);
}

void oracle_rec_box(const double *b, double *rec) {
  double vi = 1. / box_volume(b);
  rec[0] = (b[4] * b[8] - b[7] * b[5]) * vi;
  rec[4] = (b[0] * b[8] - b[6] * b[2]) * vi;
  rec[8] = (b[0] * b[4] - b[3] * b[1]) * vi;
  rec[1] = (-b[3] * b[8] + b[6] * b[5]) * vi;
  rec[2] = (b[3] * b[7] - b[6] * b[4]) * vi;
  rec[3] = (-b[1] * b[8] + b[7] * b[2]) * vi;
  rec[5] = (-b[0] * b[7] + b[6] * b[1]) * vi;
  rec[6] = (b[1] * b[5] - b[4] * b[2]) * vi;
  rec[7] = (-b[0] * b[5] + b[3] * b[2]) * vi;
}

/* fractional = rec . phys  (tensor_dot_vec) */
static void to_inter(double *ri, const double *rec, const double *rp) {
  ri[0] = rp[0] * rec[0] + rp[1] * rec[1] + rp[2] * rec[2];
  ri[1] = rp[0] * rec[3] + rp[1] * rec[4] + rp[2] * rec[5];
  ri[2] = rp[0] * rec[6] + rp[1] * rec[7] + rp[2] * rec[8];
}

/* phys = boxt^T . fractional  (tensor_t_dot_vec) */
static void to_phys(double *rp, const double *b, const double *ri) {
  rp[0] = ri[0] * b[0] + ri[1] * b[3] + ri[2] * b[6];
  rp[1] = ri[0] * b[1] + ri[1] * b[4] + ri[2] * b[7];
  rp[2] = ri[0] * b[2] + ri[1] * b[5] + ri[2] * b[8];
}

static void cprod(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* SimulationRegion::toFaceDistance (SimulationRegion_Impl.h:364-373);
 * invsqrt is 1/sqrt (source/lib/include/utilities.h). */
void oracle_face_distance(const double *b, double *dd) {
  double vol = box_volume(b), t[3];
  cprod(b + 3, b + 6, t);
  dd[0] = vol * (1. / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
  cprod(b + 6, b + 0, t);
  dd[1] = vol * (1. / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
  cprod(b + 0, b + 3, t);
  dd[2] = vol * (1. / sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
}

/* normalize_coord_cpu: source/lib/src/coord.cc:13-28 */
void oracle_normalize_coord(double *coord, int natom, const double *boxt) {
  double rec[9];
  oracle_rec_box(boxt, rec);
  for (int ii = 0; ii < natom; ++ii) {
    double ri[3];
    to_inter(ri, rec, coord + 3 * ii);
    for (int dd = 0; dd < 3; ++dd) {
      ri[dd] = fmod(ri[dd], 1.);
      if (ri[dd] < 0.) ri[dd] += 1.;
    }
    to_phys(coord + 3 * ii, boxt, ri);
  }
}

/* Cell grid of copy_coord: source/lib/src/neighbor_list.cc:756-770.
 * rc is the float cutoff promoted to double, as copy_coord_cpu does. */
void oracle_cell_grid(const double *boxt, float rcut, int *ncell, int *ngcell) {
  double to_face[3];
  double rc = (double)rcut;
  oracle_face_distance(boxt, to_face);
  for (int dd = 0; dd < 3; ++dd) {
    ncell[dd] = (int)(to_face[dd] / rc);
    if (ncell[dd] == 0) ncell[dd] = 1;
    double cell_size = to_face[dd] / ncell[dd];
    ngcell[dd] = (int)(rc / cell_size) + 1;
  }
}

/* build_clist cell index of a local atom: neighbor_list.cc:61-111 with
 * nat_stt = ext_stt = 0, nat_end = ext_end = global_grid = ncell. */
static int local_cell_index(const double *rec, const double *x, const int *ncell) {
  double inter[3];
  int idx[3];
  to_inter(inter, rec, x);
  for (int dd = 0; dd < 3; ++dd) {
    double cell_size = 1. / ncell[dd];
    idx[dd] = (int)((inter[dd] - 0.) / cell_size);
    if (inter[dd] - 0. < 0.) idx[dd]--;
    if (idx[dd] < 0) idx[dd] = 0;
    else if (idx[dd] >= ncell[dd]) idx[dd] = ncell[dd] - 1;
  }
  return (idx[0] * ncell[1] + idx[1]) * ncell[2] + idx[2];
}

/* compute_pbc_shift: neighbor_list.cc:727-743 */
static int pbc_shift(int idx, int ncell) {
  int shift = 0;
  if (idx < 0) {
    shift = 1;
    while (idx + shift * ncell < 0) shift++;
  } else if (idx >= ncell) {
    shift = -1;
    while (idx + shift * ncell >= ncell) shift--;
  }
  return shift;
}

/* Cell membership of the local atoms, in index order within each cell
 * (the push_back order of build_clist).  cell_start has nlc+1 entries. */
static void build_local_cells(const double *in_c, int nloc, const double *boxt,
                              const int *ncell, int *cell_of, int *cell_start,
                              int *cell_atoms) {
  double rec[9];
  oracle_rec_box(boxt, rec);
  int nlc = ncell[0] * ncell[1] * ncell[2];
  for (int c = 0; c <= nlc; ++c) cell_start[c] = 0;
  for (int ii = 0; ii < nloc; ++ii) {
    cell_of[ii] = local_cell_index(rec, in_c + 3 * ii, ncell);
    cell_start[cell_of[ii] + 1]++;
  }
  for (int c = 0; c < nlc; ++c) cell_start[c + 1] += cell_start[c];
  int *fill = (int *)malloc(sizeof(int) * (nlc > 0 ? nlc : 1));
  for (int c = 0; c < nlc; ++c) fill[c] = cell_start[c];
  for (int ii = 0; ii < nloc; ++ii) cell_atoms[fill[cell_of[ii]]++] = ii;
  free(fill);
}

/* copy_coord: source/lib/src/neighbor_list.cc:745-845 (called through
 * copy_coord_cpu, coord.cc:30-66).  Local atoms keep their indices; ghost
 * atoms follow, one extended cell at a time in (x, y, z) loop order, each cell
 * holding the atoms of its periodic image cell in index order.
 * Returns nall, or -(needed nall) when cap is too small. */
long oracle_copy_coord(double *out_c, int *out_t, int *mapping, long cap,
                       const double *in_c, const int *in_t, int nloc,
                       float rcut, const double *boxt) {
  int ncell[3], ngcell[3];
  oracle_cell_grid(boxt, rcut, ncell, ngcell);
  int nlc = ncell[0] * ncell[1] * ncell[2];
  int *cell_of = (int *)malloc(sizeof(int) * (nloc > 0 ? nloc : 1));
  int *cell_start = (int *)malloc(sizeof(int) * (nlc + 1));
  int *cell_atoms = (int *)malloc(sizeof(int) * (nloc > 0 ? nloc : 1));
  build_local_cells(in_c, nloc, boxt, ncell, cell_of, cell_start, cell_atoms);

  long nall = nloc;
  for (int ii0 = -ngcell[0]; ii0 < ncell[0] + ngcell[0]; ++ii0)
    for (int ii1 = -ngcell[1]; ii1 < ncell[1] + ngcell[1]; ++ii1)
      for (int ii2 = -ngcell[2]; ii2 < ncell[2] + ngcell[2]; ++ii2) {
        if (ii0 >= 0 && ii0 < ncell[0] && ii1 >= 0 && ii1 < ncell[1] &&
            ii2 >= 0 && ii2 < ncell[2])
          continue;
        int j0 = ii0 + pbc_shift(ii0, ncell[0]) * ncell[0];
        int j1 = ii1 + pbc_shift(ii1, ncell[1]) * ncell[1];
        int j2 = ii2 + pbc_shift(ii2, ncell[2]) * ncell[2];
        int c = (j0 * ncell[1] + j1) * ncell[2] + j2;
        nall += cell_start[c + 1] - cell_start[c];
      }
  if (nall > cap) {
    free(cell_of);
    free(cell_start);
    free(cell_atoms);
    return -nall;
  }
  for (int ii = 0; ii < nloc; ++ii) {
    out_c[3 * ii + 0] = in_c[3 * ii + 0];
    out_c[3 * ii + 1] = in_c[3 * ii + 1];
    out_c[3 * ii + 2] = in_c[3 * ii + 2];
    out_t[ii] = in_t[ii];
    mapping[ii] = ii;
  }
  long pos = nloc;
  for (int ii0 = -ngcell[0]; ii0 < ncell[0] + ngcell[0]; ++ii0) {
    int s0 = pbc_shift(ii0, ncell[0]);
    for (int ii1 = -ngcell[1]; ii1 < ncell[1] + ngcell[1]; ++ii1) {
      int s1 = pbc_shift(ii1, ncell[1]);
      for (int ii2 = -ngcell[2]; ii2 < ncell[2] + ngcell[2]; ++ii2) {
        int s2 = pbc_shift(ii2, ncell[2]);
        if (ii0 >= 0 && ii0 < ncell[0] && ii1 >= 0 && ii1 < ncell[1] &&
            ii2 >= 0 && ii2 < ncell[2])
          continue;
        double sd[3] = {(double)s0, (double)s1, (double)s2}, sv[3];
        to_phys(sv, boxt, sd); /* SimulationRegion::inter2Phys */
        int j0 = ii0 + s0 * ncell[0], j1 = ii1 + s1 * ncell[1],
            j2 = ii2 + s2 * ncell[2];
        int c = (j0 * ncell[1] + j1) * ncell[2] + j2;
        for (int k = cell_start[c]; k < cell_start[c + 1]; ++k) {
          int p = cell_atoms[k];
          out_c[3 * pos + 0] = in_c[3 * p + 0] - sv[0];
          out_c[3 * pos + 1] = in_c[3 * p + 1] - sv[1];
          out_c[3 * pos + 2] = in_c[3 * p + 2] - sv[2];
          out_t[pos] = in_t[p];
          mapping[pos] = p;
          pos++;
        }
      }
    }
  }
  free(cell_of);
  free(cell_start);
  free(cell_atoms);
  return nall;
}

/* build_nlist_cpu: source/lib/src/neighbor_list.cc:873-916.  Brute-force
 * candidate search over all nall atoms with the double test
 * |x_i - x_j|^2 < (float)(rcut*rcut).  cand is [nloc, mem]; returns 0 or 1
 * when some list exceeds mem (max_list_size then holds the needed size). */
int oracle_build_nlist_brute(int *numneigh, int *cand, int mem,
                             int *max_list_size, const double *c_cpy,
                             int nloc, int nall, float rcut) {
  double rcut2 = (double)(rcut * rcut);
  int overflow = 0, maxn = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : overflow) reduction(max : maxn)
  for (int ii = 0; ii < nloc; ++ii) {
    int n = 0;
    for (int jj = 0; jj < nall; ++jj) {
      if (jj == ii) continue;
      double d0 = c_cpy[ii * 3 + 0] - c_cpy[jj * 3 + 0];
      double d1 = c_cpy[ii * 3 + 1] - c_cpy[jj * 3 + 1];
      double d2 = c_cpy[ii * 3 + 2] - c_cpy[jj * 3 + 2];
      double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (r2 < rcut2) {
        if (n < mem) cand[(long)ii * mem + n] = jj;
        n++;
      }
    }
    numneigh[ii] = n;
    if (n > mem) overflow = 1;
    if (n > maxn) maxn = n;
  }
  *max_list_size = maxn;
  return overflow;
}

/* Same candidate sets as oracle_build_nlist_brute, found with a cell grid
 * over the extended system: every extended atom is binned by its own
 * fractional coordinate into cells of at least rc across (face distance), and
 * each local atom scans the cells within +-nsearch.  The formatted list does
 * not depend on candidate order, only on the candidate set. */
int oracle_build_nlist_cell(int *numneigh, int *cand, int mem,
                            int *max_list_size, const double *c_cpy, int nloc,
                            int nall, float rcut, const double *boxt) {
  double rcut2 = (double)(rcut * rcut);
  double rec[9], to_face[3];
  oracle_rec_box(boxt, rec);
  oracle_face_distance(boxt, to_face);
  int ncell[3], ngcell[3], nsearch = 1;
  oracle_cell_grid(boxt, rcut, ncell, ngcell);
  for (int dd = 0; dd < 3; ++dd)
    if (to_face[dd] / ncell[dd] - (double)rcut < 1e-6 * to_face[dd]) nsearch = 2;
  /* extended grid covers [-ngcell, ncell+ngcell) plus one guard layer */
  int lo[3], ext[3];
  for (int dd = 0; dd < 3; ++dd) {
    lo[dd] = -ngcell[dd] - 1;
    ext[dd] = ncell[dd] + 2 * ngcell[dd] + 2;
  }
  long next = (long)ext[0] * ext[1] * ext[2];
  int *cof = (int *)malloc(sizeof(int) * (nall > 0 ? nall : 1));
  int *cstart = (int *)calloc(next + 1, sizeof(int));
  int *catoms = (int *)malloc(sizeof(int) * (nall > 0 ? nall : 1));
  for (int a = 0; a < nall; ++a) {
    double f[3];
    int id[3];
    to_inter(f, rec, c_cpy + 3 * a);
    for (int dd = 0; dd < 3; ++dd) {
      id[dd] = (int)floor(f[dd] * ncell[dd]) - lo[dd];
      if (id[dd] < 0) id[dd] = 0;
      if (id[dd] >= ext[dd]) id[dd] = ext[dd] - 1;
    }
    cof[a] = (id[0] * ext[1] + id[1]) * ext[2] + id[2];
    cstart[cof[a] + 1]++;
  }
  for (long c = 0; c < next; ++c) cstart[c + 1] += cstart[c];
  int *fill = (int *)malloc(sizeof(int) * next);
  for (long c = 0; c < next; ++c) fill[c] = cstart[c];
  for (int a = 0; a < nall; ++a) catoms[fill[cof[a]]++] = a;
  free(fill);

  int overflow = 0, maxn = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(| : overflow) reduction(max : maxn)
  for (int ii = 0; ii < nloc; ++ii) {
    int c = cof[ii];
    int i2 = c % ext[2], i1 = (c / ext[2]) % ext[1], i0 = c / (ext[1] * ext[2]);
    int n = 0;
    for (int d0 = -nsearch; d0 <= nsearch; ++d0)
      for (int d1 = -nsearch; d1 <= nsearch; ++d1)
        for (int d2 = -nsearch; d2 <= nsearch; ++d2) {
          int k0 = i0 + d0, k1 = i1 + d1, k2 = i2 + d2;
          if (k0 < 0 || k1 < 0 || k2 < 0 || k0 >= ext[0] || k1 >= ext[1] ||
              k2 >= ext[2])
            continue;
          int cc = (k0 * ext[1] + k1) * ext[2] + k2;
          for (int q = cstart[cc]; q < cstart[cc + 1]; ++q) {
            int jj = catoms[q];
            if (jj == ii) continue;
            double e0 = c_cpy[ii * 3 + 0] - c_cpy[jj * 3 + 0];
            double e1 = c_cpy[ii * 3 + 1] - c_cpy[jj * 3 + 1];
            double e2 = c_cpy[ii * 3 + 2] - c_cpy[jj * 3 + 2];
            double r2 = e0 * e0 + e1 * e1 + e2 * e2;
            if (r2 < rcut2) {
              if (n < mem) cand[(long)ii * mem + n] = jj;
              n++;
            }
          }
        }
    numneigh[ii] = n;
    if (n > mem) overflow = 1;
    if (n > maxn) maxn = n;
  }
  *max_list_size = maxn;
  free(cof);
  free(cstart);
  free(catoms);
  return overflow;
}

/* format_nlist_i_cpu: source/lib/src/fmt_nlist.cc:98-143.  Keys are
 * (type, float r^2 from float-rounded coordinates, index); neighbors beyond
 * sel of their type are dropped (the nearest are kept).  Returns -1, or the
 * last overflowing type. */
typedef struct {
  int type;
  float dist;
  int index;
} nbr_key;

static int key_less(const nbr_key *a, const nbr_key *b) {
  if (a->type != b->type) return a->type < b->type;
  if (a->dist != b->dist) return a->dist < b->dist;
  return a->index < b->index;
}

static void sort_keys(nbr_key *v, int n) {
  /* insertion sort: candidate lists are short (a few hundred at most) */
  for (int i = 1; i < n; ++i) {
    nbr_key x = v[i];
    int j = i - 1;
    while (j >= 0 && key_less(&x, &v[j])) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
}

int oracle_format_nlist_i(int *fmt, const double *posi, const int *type,
                          int i_idx, const int *cand, int ncand, float rcut,
                          const int *sec, int ntypes, nbr_key *work) {
  int nnei = sec[ntypes];
  for (int k = 0; k < nnei; ++k) fmt[k] = -1;
  float rcut2 = rcut * rcut;
  int n = 0;
  for (int kk = 0; kk < ncand; ++kk) {
    int j = cand[kk];
    if (type[j] < 0) continue;
    float d0 = (float)posi[j * 3 + 0] - (float)posi[i_idx * 3 + 0];
    float d1 = (float)posi[j * 3 + 1] - (float)posi[i_idx * 3 + 1];
    float d2 = (float)posi[j * 3 + 2] - (float)posi[i_idx * 3 + 2];
    float rr2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (rr2 <= rcut2) {
      work[n].type = type[j];
      work[n].dist = rr2;
      work[n].index = j;
      n++;
    }
  }
  sort_keys(work, n);
  int iter[64];
  for (int t = 0; t < ntypes; ++t) iter[t] = sec[t];
  int overflowed = -1;
  for (int kk = 0; kk < n; ++kk) {
    int t = work[kk].type;
    if (iter[t] < sec[t + 1])
      fmt[iter[t]++] = work[kk].index;
    else
      overflowed = t;
  }
  return overflowed;
}

/* format_nlist_cpu over all local atoms (fmt_nlist.cc:145-177).  overflow[i]
 * receives format_nlist_i's return code. */
void oracle_format_nlist(int *nlist, int *overflow, const double *posi,
                         const int *type, int nloc, const int *numneigh,
                         const int *cand, int mem, float rcut, const int *sec,
                         int ntypes) {
#pragma omp parallel
  {
    nbr_key *work = (nbr_key *)malloc(sizeof(nbr_key) * (mem > 0 ? mem : 1));
#pragma omp for schedule(dynamic, 16)
    for (int ii = 0; ii < nloc; ++ii) {
      int r = oracle_format_nlist_i(nlist + (long)ii * sec[ntypes], posi, type,
                                    ii, cand + (long)ii * mem, numneigh[ii],
                                    rcut, sec, ntypes, work);
      if (overflow) overflow[ii] = r;
    }
    free(work);
  }
}

/* spline5_switch: source/lib/include/switcher.h:60-85 */
static void spline5(double *vv, double *dd, double xx, float rmin, float rmax) {
  if (xx < rmin) {
    *dd = 0.;
    *vv = 1.;
  } else if (xx < rmax) {
    double uu = (xx - rmin) / (float)(rmax - rmin);
    double du = 1. / (float)(rmax - rmin);
    *vv = uu * uu * uu * (-6. * uu * uu + 15. * uu - 10.) + 1.;
    *dd = (3. * uu * uu * (-6. * uu * uu + 15. * uu - 10.) +
           uu * uu * uu * (-12. * uu + 15.)) *
          du;
  } else {
    *dd = 0.;
    *vv = 0.;
  }
}

/* env_mat_a_cpu (source/lib/src/env_mat.cc:123-221) followed by the
 * davg/dstd normalisation of prod_env_mat_a_cpu (prod_env_mat.cc:66-107) for
 * one centre atom.  em: [nnei*4], em_deriv: [nnei*12], rij: [nnei*3]. */
static void env_mat_a_one(double *em, double *em_deriv, double *rij,
                          const double *posi, int ti, int i_idx,
                          const int *fmt, const int *sec, int ntypes,
                          float rmin, float rmax, const double *avg,
                          const double *std) {
  int nnei = sec[ntypes];
  int nem = nnei * 4;
  double *dem = (double *)calloc(nem, sizeof(double));
  double *dder = (double *)calloc(nem * 3, sizeof(double));
  for (int k = 0; k < nnei * 3; ++k) rij[k] = 0.;
  for (int s = 0; s < ntypes; ++s) {
    for (int jj = sec[s]; jj < sec[s + 1]; ++jj) {
      if (fmt[jj] < 0) break;
      int j = fmt[jj];
      for (int dd = 0; dd < 3; ++dd)
        rij[jj * 3 + dd] = posi[j * 3 + dd] - posi[i_idx * 3 + dd];
    }
  }
  for (int s = 0; s < ntypes; ++s) {
    for (int jj = sec[s]; jj < sec[s + 1]; ++jj) {
      if (fmt[jj] < 0) break;
      const double *rr = rij + jj * 3;
      double nr2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
      double inr = 1. / sqrt(nr2);
      double nr = nr2 * inr;
      double inr2 = inr * inr;
      double inr4 = inr2 * inr2;
      double inr3 = inr4 * nr;
      double sw, dsw;
      spline5(&sw, &dsw, nr, rmin, rmax);
      double *v = dem + jj * 4;
      double *d = dder + jj * 12;
      v[0] = 1. / nr;
      v[1] = rr[0] / nr2;
      v[2] = rr[1] / nr2;
      v[3] = rr[2] / nr2;
      for (int dd = 0; dd < 3; ++dd)
        d[dd] = rr[dd] * inr3 * sw - v[0] * dsw * rr[dd] * inr;
      for (int c = 0; c < 3; ++c)
        for (int dd = 0; dd < 3; ++dd) {
          double t = 2. * rr[c] * rr[dd] * inr4;
          if (c == dd) t = t - inr2;
          d[3 + 3 * c + dd] = t * sw - v[1 + c] * dsw * rr[dd] * inr;
        }
      v[0] *= sw;
      v[1] *= sw;
      v[2] *= sw;
      v[3] *= sw;
    }
  }
  for (int k = 0; k < nem; ++k) {
    if (ti >= 0)
      em[k] = (dem[k] - avg[ti * nem + k]) / std[ti * nem + k];
    else
      em[k] = 0.;
  }
  for (int k = 0; k < nem * 3; ++k) {
    if (ti >= 0)
      em_deriv[k] = dder[k] / std[ti * nem + k / 3];
    else
      em_deriv[k] = 0.;
  }
  free(dem);
  free(dder);
}

/* prod_env_mat_a_cpu on an already formatted nlist (prod_env_mat.cc:14-108
 * minus the per-atom format step, which oracle_format_nlist performs). */
void oracle_env_mat_a(double *em, double *em_deriv, double *rij,
                      const int *nlist, const double *posi, const int *type,
                      int nloc, const double *avg, const double *std,
                      float rcut, float rcut_smth, const int *sec, int ntypes) {
  int nnei = sec[ntypes];
#pragma omp parallel for schedule(static)
  for (int ii = 0; ii < nloc; ++ii) {
    env_mat_a_one(em + (long)ii * nnei * 4, em_deriv + (long)ii * nnei * 12,
                  rij + (long)ii * nnei * 3, posi, type[ii], ii,
                  nlist + (long)ii * nnei, sec, ntypes, rcut_smth, rcut, avg,
                  std);
  }
}

/* locate_xx: source/lib/src/tabulate.cc:19-42 */
static void locate_xx(double lower, double upper, double mx, double s0,
                      double s1, double *xx, int *idx) {
  if (*xx < lower) {
    *idx = 0;
    *xx = 0.;
  } else if (*xx < upper) {
    *idx = (int)((*xx - lower) / s0);
    *xx -= (*idx * s0 + lower);
  } else if (*xx < mx) {
    int first = (int)((upper - lower) / s0);
    *idx = first + (int)((*xx - upper) / s1);
    *xx -= ((*idx - first) * s1 + upper);
  } else {
    *idx = (int)((upper - lower) / s0) + (int)((mx - upper) / s1) - 1;
    *xx = 0.;
  }
}

/* tabulate_fusion_se_a_cpu: source/lib/src/tabulate.cc:80-156 (no
 * two_embed).  out: [nloc, 4, M]. */
void oracle_tabulate_fusion_se_a(double *out, const double *table,
                                 const double *info, const double *em_x,
                                 const double *em, int nloc, int nnei, int M,
                                 int is_sorted) {
  memset(out, 0, sizeof(double) * (size_t)nloc * 4 * M);
  const double lower = info[0], upper = info[1], mx = info[2], s0 = info[3],
               s1 = info[4];
#pragma omp parallel for schedule(static)
  for (int ii = 0; ii < nloc; ii++) {
    double ago = em_x[(long)ii * nnei + nnei - 1];
    int unloop = 0;
    double *o = out + (long)ii * M * 4;
    for (int jj = 0; jj < nnei; jj++) {
      const double *ll = em + (long)ii * nnei * 4 + jj * 4;
      double xx = em_x[(long)ii * nnei + jj];
      if (ago == xx && ll[1] == 0. && ll[2] == 0. && ll[3] == 0. && is_sorted)
        unloop = 1;
      int idx = 0;
      locate_xx(lower, upper, mx, s0, s1, &xx, &idx);
      for (int kk = 0; kk < M; kk++) {
        const double *a = table + (long)idx * M * 6 + 6 * kk;
        double var = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * xx) * xx) * xx) * xx) * xx;
        if (unloop) {
          o[0 * M + kk] += (nnei - jj) * var * ll[0];
          o[1 * M + kk] += (nnei - jj) * var * ll[1];
          o[2 * M + kk] += (nnei - jj) * var * ll[2];
          o[3 * M + kk] += (nnei - jj) * var * ll[3];
        } else {
          o[0 * M + kk] += var * ll[0];
          o[1 * M + kk] += var * ll[1];
          o[2 * M + kk] += var * ll[2];
          o[3 * M + kk] += var * ll[3];
        }
      }
      if (unloop) break;
    }
  }
}

/* tabulate_fusion_se_a_grad_cpu: source/lib/src/tabulate.cc:158-259 */
void oracle_tabulate_fusion_se_a_grad(double *dy_dem_x, double *dy_dem,
                                      const double *table, const double *info,
                                      const double *em_x, const double *em,
                                      const double *dy, int nloc, int nnei,
                                      int M, int is_sorted) {
  memset(dy_dem_x, 0, sizeof(double) * (size_t)nloc * nnei);
  memset(dy_dem, 0, sizeof(double) * (size_t)nloc * nnei * 4);
  const double lower = info[0], upper = info[1], mx = info[2], s0 = info[3],
               s1 = info[4];
#pragma omp parallel for schedule(static)
  for (int ii = 0; ii < nloc; ii++) {
    double ago = em_x[(long)ii * nnei + nnei - 1];
    int unloop = 0;
    for (int jj = 0; jj < nnei; jj++) {
      const double *ll = em + (long)ii * nnei * 4 + jj * 4;
      double *dd = dy_dem + (long)ii * nnei * 4 + jj * 4;
      double xx = em_x[(long)ii * nnei + jj];
      if (ago == xx && ll[1] == 0. && ll[2] == 0. && ll[3] == 0. && is_sorted)
        unloop = 1;
      int idx = 0;
      locate_xx(lower, upper, mx, s0, s1, &xx, &idx);
      double grad = 0.;
      for (int kk = 0; kk < M; kk++) {
        const double *rr = dy + (long)ii * M * 4;
        double r0 = rr[0 * M + kk], r1 = rr[1 * M + kk], r2 = rr[2 * M + kk],
               r3 = rr[3 * M + kk];
        const double *a = table + (long)idx * M * 6 + 6 * kk;
        double res = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * xx) * xx) * xx) * xx) * xx;
        double g = (a[1] + (2 * a[2] + (3 * a[3] + (4 * a[4] + 5 * a[5] * xx) * xx) * xx) * xx);
        double dot = ll[0] * r0 + ll[1] * r1 + ll[2] * r2 + ll[3] * r3;
        if (unloop) {
          grad += g * dot * (nnei - jj);
          dd[0] += res * r0 * (nnei - jj);
          dd[1] += res * r1 * (nnei - jj);
          dd[2] += res * r2 * (nnei - jj);
          dd[3] += res * r3 * (nnei - jj);
        } else {
          grad += g * dot;
          dd[0] += res * r0;
          dd[1] += res * r1;
          dd[2] += res * r2;
          dd[3] += res * r3;
        }
      }
      dy_dem_x[(long)ii * nnei + jj] = grad;
      if (unloop) break;
    }
  }
}

/* tabulate_fusion_se_a_grad_grad_cpu: source/lib/src/tabulate.cc:261-383 */
void oracle_tabulate_fusion_se_a_grad_grad(
    double *dz_dy, const double *table, const double *info, const double *em_x,
    const double *em, const double *dz_dy_dem_x, const double *dz_dy_dem,
    int nloc, int nnei, int M, int is_sorted) {
  memset(dz_dy, 0, sizeof(double) * (size_t)nloc * 4 * M);
  const double lower = info[0], upper = info[1], mx = info[2], s0 = info[3],
               s1 = info[4];
#pragma omp parallel for schedule(static)
  for (int ii = 0; ii < nloc; ii++) {
    double ago = em_x[(long)ii * nnei + nnei - 1];
    int unloop = 0;
    double *o = dz_dy + (long)ii * M * 4;
    for (int jj = 0; jj < nnei; jj++) {
      const double *ll = em + (long)ii * nnei * 4 + jj * 4;
      const double *hh = dz_dy_dem + (long)ii * nnei * 4 + jj * 4;
      double xx = em_x[(long)ii * nnei + jj];
      double dz_xx = dz_dy_dem_x[(long)ii * nnei + jj];
      if (ago == xx && ll[1] == 0. && ll[2] == 0. && ll[3] == 0. && is_sorted)
        unloop = 1;
      int idx = 0;
      locate_xx(lower, upper, mx, s0, s1, &xx, &idx);
      for (int kk = 0; kk < M; kk++) {
        const double *a = table + (long)idx * M * 6 + 6 * kk;
        double var = a[0] + (a[1] + (a[2] + (a[3] + (a[4] + a[5] * xx) * xx) * xx) * xx) * xx;
        double vg = a[1] + (2. * a[2] + (3. * a[3] + (4. * a[4] + 5. * a[5] * xx) * xx) * xx) * xx;
        double two_grad = 0.;
        for (int c = 0; c < 4; ++c) {
          double t = var * hh[c] + (dz_xx * vg + two_grad) * ll[c];
          if (unloop)
            o[c * M + kk] += (nnei - jj) * t;
          else
            o[c * M + kk] += t;
        }
      }
      if (unloop) break;
    }
  }
}

/* prod_force_a_cpu: source/lib/src/prod_force.cc:23-76 (all frames, single
 * thread range).  force: [nframes, nall, 3]. */
void oracle_prod_force_a(double *force, const double *net_deriv,
                         const double *env_deriv, const int *nlist, int nloc,
                         int nall, int nnei, int nframes) {
  const int ndescrpt = 4 * nnei;
  memset(force, 0, sizeof(double) * (size_t)nframes * nall * 3);
  for (long i_idx = 0; i_idx < (long)nframes * nloc; ++i_idx) {
    long kk = i_idx / nloc, ll = i_idx % nloc;
    long i_nall = kk * nall + ll;
    for (int aa = 0; aa < ndescrpt; ++aa) {
      double nd = net_deriv[i_idx * ndescrpt + aa];
      const double *ed = env_deriv + i_idx * ndescrpt * 3 + aa * 3;
      force[i_nall * 3 + 0] -= nd * ed[0];
      force[i_nall * 3 + 1] -= nd * ed[1];
      force[i_nall * 3 + 2] -= nd * ed[2];
    }
    for (int jj = 0; jj < nnei; ++jj) {
      int j_idx = nlist[i_idx * nnei + jj];
      if (j_idx < 0) continue;
      for (int aa = jj * 4; aa < jj * 4 + 4; ++aa) {
        double nd = net_deriv[i_idx * ndescrpt + aa];
        const double *ed = env_deriv + i_idx * ndescrpt * 3 + aa * 3;
        force[kk * nall * 3 + (long)j_idx * 3 + 0] += nd * ed[0];
        force[kk * nall * 3 + (long)j_idx * 3 + 1] += nd * ed[1];
        force[kk * nall * 3 + (long)j_idx * 3 + 2] += nd * ed[2];
      }
    }
  }
}

/* prod_virial_a_cpu: source/lib/src/prod_virial.cc:22-70 (one frame,
 * sequential accumulation order).  virial: [9], atom_virial: [nall, 9]. */
void oracle_prod_virial_a(double *virial, double *atom_virial,
                          const double *net_deriv, const double *env_deriv,
                          const double *rij, const int *nlist, int nloc,
                          int nall, int nnei) {
  const int ndescrpt = 4 * nnei;
  for (int k = 0; k < 9; ++k) virial[k] = 0.;
  memset(atom_virial, 0, sizeof(double) * (size_t)nall * 9);
  for (long ii = 0; ii < nloc; ++ii) {
    for (int jj = 0; jj < nnei; ++jj) {
      int j_idx = nlist[ii * nnei + jj];
      if (j_idx < 0) continue;
      for (int aa = jj * 4; aa < jj * 4 + 4; ++aa) {
        double pref = -1.0 * net_deriv[ii * ndescrpt + aa];
        for (int d0 = 0; d0 < 3; ++d0)
          for (int d1 = 0; d1 < 3; ++d1) {
            double tmp = pref * rij[ii * nnei * 3 + jj * 3 + d1] *
                         env_deriv[ii * ndescrpt * 3 + aa * 3 + d0];
            virial[d0 * 3 + d1] -= tmp;
            atom_virial[(long)j_idx * 9 + d0 * 3 + d1] -= tmp;
          }
      }
    }
  }
}

/* prod_force_grad_a_cpu: source/lib/src/prod_force_grad.cc:22-79 (the
 * transpose of prod_force_a with the neighbor's force gradient read from the
 * frame's local atoms, j -> j % nloc).  grad: [nframes, nloc, 3];
 * grad_net: [nframes, nloc, 4 nnei]. */
void oracle_prod_force_grad_a(double *grad_net, const double *grad,
                              const double *env_deriv, const int *nlist,
                              int nloc, int nnei, int nframes) {
  const int ndescrpt = 4 * nnei;
  for (long ii = 0; ii < (long)nframes * nloc; ++ii) {
    const long kk = ii / nloc;
    double *gn = grad_net + ii * ndescrpt;
    const double *gi = grad + ii * 3;
    for (int aa = 0; aa < ndescrpt; ++aa) {
      const double *ed = env_deriv + ii * ndescrpt * 3 + aa * 3;
      double acc = 0.;
      for (int dd = 0; dd < 3; ++dd) acc -= gi[dd] * ed[dd];
      gn[aa] = acc;
    }
    for (int jj = 0; jj < nnei; ++jj) {
      int j_idx = nlist[ii * nnei + jj];
      if (j_idx >= nloc) j_idx = j_idx % nloc;
      if (j_idx < 0) continue;
      const double *gj = grad + (kk * nloc + j_idx) * 3;
      for (int aa = jj * 4; aa < jj * 4 + 4; ++aa) {
        const double *ed = env_deriv + ii * ndescrpt * 3 + aa * 3;
        for (int dd = 0; dd < 3; ++dd) gn[aa] += gj[dd] * ed[dd];
      }
    }
  }
}

/* prod_virial_grad_a_cpu: source/lib/src/prod_virial_grad.cc:20-62 (one
 * frame): grad_net[i, 4j+a] = sum_{d0,d1} grad[d0,d1] rij[i,j,d1]
 * env_deriv[i,4j+a,d0] over valid neighbors, 0 elsewhere. */
void oracle_prod_virial_grad_a(double *grad_net, const double *grad,
                               const double *env_deriv, const double *rij,
                               const int *nlist, int nloc, int nnei) {
  const int ndescrpt = 4 * nnei;
  memset(grad_net, 0, sizeof(double) * (size_t)nloc * ndescrpt);
  for (long ii = 0; ii < nloc; ++ii) {
    for (int jj = 0; jj < nnei; ++jj) {
      if (nlist[ii * nnei + jj] < 0) continue;
      const double *r = rij + ii * nnei * 3 + jj * 3;
      for (int aa = jj * 4; aa < jj * 4 + 4; ++aa) {
        const double *ed = env_deriv + ii * ndescrpt * 3 + aa * 3;
        for (int d0 = 0; d0 < 3; ++d0)
          for (int d1 = 0; d1 < 3; ++d1)
            grad_net[ii * ndescrpt + aa] -= -1.0 * grad[d0 * 3 + d1] * r[d1] * ed[d0];
      }
    }
  }
}

/* Number of OpenMP threads the oracle runs with (1 without OpenMP). */
int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
