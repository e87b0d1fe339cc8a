// C-ABI harness around the reference's own CPU library functions.
//
// TEST INFRASTRUCTURE ONLY.  This file contains no reference code: it only
// calls the reference functions (compiled from their sources in place, see
// oracle/Makefile) through extern "C" entry points so that the oracle and the
// HIP kernels can be compared with the real reference on the same inputs.
// The resulting library lives in oracle/_ref/ and is never linked by the
// product.
#include <cstdint>
#include <vector>

#include "coord.h"
#include "fmt_nlist.h"
#include "neighbor_list.h"
#include "prod_env_mat.h"
#include "prod_force.h"
#include "prod_force_grad.h"
#include "prod_virial.h"
#include "prod_virial_grad.h"
#include "region.h"
#include "tabulate.h"

extern "C" {

// deepmd::normalize_coord_cpu (source/lib/src/coord.cc:13-28)
void ref_normalize_coord(double* coord, int natom, const double* box) {
  deepmd::Region<double> region;
  deepmd::init_region_cpu(region, box);
  deepmd::normalize_coord_cpu(coord, natom, region);
}

// deepmd::copy_coord_cpu (source/lib/src/coord.cc:30-66); returns 0 or 1
int ref_copy_coord(double* out_c, int* out_t, int* mapping, int* nall,
                   const double* in_c, const int* in_t, int nloc, int mem_nall,
                   float rcut, const double* box) {
  deepmd::Region<double> region;
  deepmd::init_region_cpu(region, box);
  return deepmd::copy_coord_cpu(out_c, out_t, mapping, nall, in_c, in_t, nloc,
                                mem_nall, rcut, region);
}

// deepmd::build_nlist_cpu (source/lib/src/neighbor_list.cc:873-916)
// cand is [nloc, mem]
int ref_build_nlist(int* numneigh, int* cand, int mem, int* max_list_size,
                    const double* c_cpy, int nloc, int nall, float rcut) {
  std::vector<int> ilist(nloc);
  std::vector<int*> firstneigh(nloc);
  for (int ii = 0; ii < nloc; ++ii) firstneigh[ii] = cand + (int64_t)ii * mem;
  deepmd::InputNlist inlist(nloc, ilist.data(), numneigh, firstneigh.data());
  return deepmd::build_nlist_cpu(inlist, max_list_size, c_cpy, nloc, nall, mem,
                                 rcut);
}

// deepmd::prod_env_mat_a_cpu (source/lib/src/prod_env_mat.cc:14-108);
// candidate lists come from ref_build_nlist.
void ref_prod_env_mat_a(double* em, double* em_deriv, double* rij, int* nlist,
                        const double* coord, const int* type, int nloc,
                        int nall, const int* numneigh, const int* cand,
                        int mem, const double* avg, const double* std,
                        float rcut, float rcut_smth, const int* sel,
                        int ntypes) {
  std::vector<int> ilist(nloc), nn(numneigh, numneigh + nloc);
  std::vector<int*> firstneigh(nloc);
  for (int ii = 0; ii < nloc; ++ii) {
    ilist[ii] = ii;
    firstneigh[ii] = const_cast<int*>(cand) + (int64_t)ii * mem;
  }
  deepmd::InputNlist inlist(nloc, ilist.data(), nn.data(), firstneigh.data());
  std::vector<int> sec(ntypes + 1, 0);
  for (int t = 0; t < ntypes; ++t) sec[t + 1] = sec[t] + sel[t];
  deepmd::prod_env_mat_a_cpu(em, em_deriv, rij, nlist, coord, type, inlist,
                             mem, avg, std, nloc, nall, rcut, rcut_smth, sec);
}

// deepmd::tabulate_fusion_se_a_cpu (source/lib/src/tabulate.cc:80-156)
void ref_tabulate_fusion_se_a(double* out, const double* table,
                              const double* info, const double* em_x,
                              const double* em, int nloc, int nnei, int M) {
  deepmd::tabulate_fusion_se_a_cpu(out, table, info, em_x, em,
                                   (const double*)nullptr, nloc, nnei, M);
}

void ref_tabulate_fusion_se_a_f32(float* out, const float* table,
                                  const float* info, const float* em_x,
                                  const float* em, int nloc, int nnei, int M) {
  deepmd::tabulate_fusion_se_a_cpu(out, table, info, em_x, em,
                                   (const float*)nullptr, nloc, nnei, M);
}

// deepmd::tabulate_fusion_se_a_grad_cpu (tabulate.cc:158-259)
void ref_tabulate_fusion_se_a_grad(double* dy_dem_x, double* dy_dem,
                                   const double* table, const double* info,
                                   const double* em_x, const double* em,
                                   const double* dy, int nloc, int nnei,
                                   int M) {
  deepmd::tabulate_fusion_se_a_grad_cpu(dy_dem_x, dy_dem, (double*)nullptr,
                                        table, info, em_x, em,
                                        (const double*)nullptr, dy, nloc, nnei,
                                        M);
}

// deepmd::tabulate_fusion_se_a_grad_grad_cpu (tabulate.cc:261-383)
void ref_tabulate_fusion_se_a_grad_grad(double* dz_dy, const double* table,
                                        const double* info, const double* em_x,
                                        const double* em,
                                        const double* dz_dy_dem_x,
                                        const double* dz_dy_dem, int nloc,
                                        int nnei, int M) {
  deepmd::tabulate_fusion_se_a_grad_grad_cpu(
      dz_dy, table, info, em_x, em, (const double*)nullptr, dz_dy_dem_x,
      dz_dy_dem, (const double*)nullptr, nloc, nnei, M, true);
}

// deepmd::prod_force_a_cpu (source/lib/src/prod_force.cc:23-76)
void ref_prod_force_a(double* force, const double* net_deriv,
                      const double* env_deriv, const int* nlist, int nloc,
                      int nall, int nnei, int nframes) {
  deepmd::prod_force_a_cpu(force, net_deriv, env_deriv, nlist, nloc, nall,
                           nnei, nframes);
}

// deepmd::prod_virial_a_cpu (source/lib/src/prod_virial.cc:22-70)
void ref_prod_virial_a(double* virial, double* atom_virial,
                       const double* net_deriv, const double* env_deriv,
                       const double* rij, const int* nlist, int nloc, int nall,
                       int nnei) {
  deepmd::prod_virial_a_cpu(virial, atom_virial, net_deriv, env_deriv, rij,
                            nlist, nloc, nall, nnei);
}

// deepmd::prod_force_grad_a_cpu (source/lib/src/prod_force_grad.cc:22-79)
void ref_prod_force_grad_a(double* grad_net, const double* grad,
                           const double* env_deriv, const int* nlist, int nloc,
                           int nnei, int nframes) {
  deepmd::prod_force_grad_a_cpu(grad_net, grad, env_deriv, nlist, nloc, nnei,
                                nframes);
}

// deepmd::prod_virial_grad_a_cpu (source/lib/src/prod_virial_grad.cc:20-62)
void ref_prod_virial_grad_a(double* grad_net, const double* grad,
                            const double* env_deriv, const double* rij,
                            const int* nlist, int nloc, int nnei) {
  deepmd::prod_virial_grad_a_cpu(grad_net, grad, env_deriv, rij, nlist, nloc,
                                 nnei);
}
}
