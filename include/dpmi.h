/*
 * libdpmi — MI355X (gfx950) kernels for the DeepPot-SE se_e2_a
 * descriptor-to-force path, behind a plain C ABI.
 *
 * Conventions
 *  - Every pointer argument is DEVICE memory unless its name ends in _host.
 *  - The caller allocates every output and every workspace; the library
 *    never allocates or frees device memory.
 *  - Sizes are int64_t; `stream` is a hipStream_t passed as void* (NULL is the
 *    null stream).  Work is queued on `stream`; functions return without
 *    waiting unless documented otherwise.
 *  - Return value: DPMI_OK (0) or an error code; dpmi_last_error() gives the
 *    message of the last failure on the calling thread.
 *  - Layouts are contiguous row-major, as in the reference op contracts.
 *
 * Each entry point names the reference function it replaces (paths relative
 * to the reference repository root).
 */
#ifndef DPMI_H_
#define DPMI_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPMI_OK 0
#define DPMI_ERR_ARG 1      /* invalid argument (std::invalid_argument in the reference) */
#define DPMI_ERR_HIP 2      /* HIP runtime error (deepmd_exception from DPErrcheck) */
#define DPMI_ERR_CAPACITY 3 /* output capacity too small; the needed size is returned */

const char* dpmi_last_error(void);
int dpmi_abi_version(void);

/* ------------------------------------------------------------------------
 * Neighbor list.  Replaces, as one pipeline, normalize_coord_gpu,
 * copy_coord_gpu and build_nlist_gpu (source/lib/include/coord.h:36-80,
 * neighbor_list.h:145-240) plus format_nbor_list_gpu
 * (source/lib/include/fmt_nlist.h:20-32).  The output is the formatted
 * list of format_nlist_cpu (source/lib/src/fmt_nlist.cc:98-177): per type
 * section, neighbors sorted by (float r^2, extended index), -1 padding.
 * ------------------------------------------------------------------------ */

/* Workspace bytes for dpmi_copy_coord_f64 + dpmi_build_nlist_f64.
 * box_host: 9 doubles (rows are the cell vectors), or NULL for an open
 * (non-periodic) system. */
int dpmi_nlist_workspace_bytes(int64_t* bytes, const double* box_host, int64_t nloc, float rcut);

/* Normalize local coordinates into the box and append periodic ghost atoms
 * in the layout of copy_coord_cpu (source/lib/src/coord.cc:30-66,
 * neighbor_list.cc:745-845): locals first, then ghost cells in (x, y, z) loop
 * order.  Also bins atoms into the cell grid kept in the workspace for
 * dpmi_build_nlist_f64.  Waits for the stream once to learn nall.
 * If nall > mem_nall, returns DPMI_ERR_CAPACITY with *nall set (the caller
 * grows the buffers and calls again, as _norm_copy_coord_cpu does).
 * With box_host == NULL nothing is copied and nall = nloc (the open-system
 * grid waits once more, for the bounding box of the coordinates). */
int dpmi_copy_coord_f64(double* out_coord, int* out_type, int* mapping, int64_t* nall, int64_t mem_nall,
                        const double* in_coord, const int* in_type, int64_t nloc,
                        const double* box_host, float rcut, void* workspace, int64_t workspace_bytes,
                        void* stream);

/* Formatted neighbor list [nloc, sum(sel)] in extended indices.  Candidates
 * are |x_i - x_j|^2 < rcut^2 in double (build_nlist_cpu,
 * neighbor_list.cc:873-916); kept when the float r^2 of float-rounded
 * coordinates is <= rcut^2 (fmt_nlist.cc:114-128).  overflow[i] (optional)
 * receives -1 or the largest type whose neighbors exceeded sel.
 * nlist_mapped (optional) receives mapping[nlist] (the local-index list
 * that ProdEnvMatAOp returns, source/op/tf/prod_env_mat_multi_device.cc:
 * 619-622).  cand_slot (optional, [nloc, sum(sel)] int16) receives each
 * neighbor's position in the candidate table of its centre's cell (-1 where
 * none), which dpmi_force_virial_cells_* uses; the workspace then keeps the
 * cell grid for that call.  Waits for the stream to read its status back
 * (twice when some atoms need the per-atom pass); the _nowait form below
 * does not. */
int dpmi_build_nlist_f64(int* nlist, int* overflow, int* nlist_mapped, int16_t* cand_slot,
                         const double* ext_coord, const int* ext_type, const int* mapping, int64_t nloc,
                         int64_t nall, float rcut, const int* sel_host, int ntypes, const void* workspace,
                         int64_t workspace_bytes, void* stream);

/* The same pipeline without a host wait, for periodic boxes (the form a HIP
 * graph can capture; same neighbor lists).  dpmi_copy_coord_capacity gives a
 * bound on nall from the cell grid alone (every local cell has at most
 * prod_d ceil((ncell_d + 2 ngcell_d) / ncell_d) images among the extended
 * cells); the output arrays of dpmi_copy_coord_nowait_f64 hold mem_nall >=
 * that bound entries, of which the first nall are written (nall stays in the
 * workspace).  dpmi_build_nlist_nowait_f64 takes nall from the workspace and
 * runs its per-atom pass on a fixed grid.  Neither reads anything back:
 * errors (more than 1024 candidates for an atom, a workspace of another
 * system, more ghosts than mem_nall) are bits of the status word that
 * dpmi_nlist_status copies into device memory (0 = no error; the caller
 * reads it when it chooses), and dpmi_force_virial_cells_* writes NaN forces
 * for a list whose status is not 0. */
int dpmi_copy_coord_capacity(int64_t* capacity, const double* box_host, int64_t nloc, float rcut);
int dpmi_copy_coord_nowait_f64(double* out_coord, int* out_type, int* mapping, int64_t mem_nall,
                               const double* in_coord, const int* in_type, int64_t nloc, const double* box_host,
                               float rcut, void* workspace, int64_t workspace_bytes, void* stream);
int dpmi_build_nlist_nowait_f64(int* nlist, int* overflow, int* nlist_mapped, int16_t* cand_slot,
                                const double* ext_coord, const int* ext_type, const int* mapping, int64_t nloc,
                                float rcut, const int* sel_host, int ntypes, const void* workspace,
                                int64_t workspace_bytes, void* stream);
int dpmi_nlist_status(int* status, const void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Environment matrix.  Replaces prod_env_mat_a_gpu
 * (source/lib/include/prod_env_mat.h:45-62) on an already formatted list
 * (env_mat_a_cpu, source/lib/src/env_mat.cc:123-221, and the davg/dstd
 * normalisation of prod_env_mat.cc:66-107).  em [nloc, nnei, 4],
 * em_deriv [nloc, nnei, 4, 3] (derivative w.r.t. the centre atom, already
 * divided by std; may be NULL), rij [nloc, nnei, 3] (may be NULL).
 * avg/std: [ntypes, nnei, 4].
 * ------------------------------------------------------------------------ */
int dpmi_prod_env_mat_a_f64(double* em, double* em_deriv, double* rij, const int* nlist,
                            const double* ext_coord, const int* ext_type, int64_t nloc, const double* avg,
                            const double* std, float rcut, float rcut_smth, const int* sel_host, int ntypes,
                            void* stream);
int dpmi_prod_env_mat_a_f32(float* em, float* em_deriv, float* rij, const int* nlist,
                            const float* ext_coord, const int* ext_type, int64_t nloc, const float* avg,
                            const float* std, float rcut, float rcut_smth, const int* sel_host, int ntypes,
                            void* stream);
/* Precision form: double coordinates, the difference x_j - x_i formed in
 * double and everything after it in float.  (Both float forms evaluate with
 * hardware reciprocals - 1/r from rsq, 1/(rcut - rcut_smth) once, times
 * 1/std - a few ulp from the IEEE divisions of the double form.) */
int dpmi_prod_env_mat_a_f32c64(float* em, float* em_deriv, float* rij, const int* nlist, const double* ext_coord,
                               const int* ext_type, int64_t nloc, const float* avg, const float* std, float rcut,
                               float rcut_smth, const int* sel_host, int ntypes, void* stream);

/* ------------------------------------------------------------------------
 * Environment matrix with the PyTorch backend's semantics: replaces
 * prod_env_mat / _make_env_mat (deepmd/pt/model/descriptor/env_mat.py:11-92,
 * smooth weights deepmd/pt/utils/preprocess.py:9-34), the function
 * DescrptBlockSeA.forward calls (deepmd/pt/model/descriptor/se_a.py:781-790),
 * and the two products its autograd forms.  Any list: nlist [nframes, nloc,
 * nnei] of extended indices in any order, -1 padding anywhere, neighbors
 * beyond rcut allowed (smooth weight 0).  ext_coord [nframes, nall, 3];
 * atype [nframes, nloc] in [0, ntypes); mean / stddev [ntypes, nnei, ncomp]
 * with ncomp = 4, or 1 for radial_only.  Per slot, d = x_j - x_i, r = |d|,
 * q = r + protection, w = compute_smooth_weight(r, rcut_smth, rcut) (or
 * compute_exp_sw with use_exp_switch):
 *   env_mat [nframes, nloc, nnei, ncomp] = ([1/q, d/q^2] w - mean) / stddev
 *   diff [nframes, nloc, nnei, 3] = d, sw [nframes, nloc, nnei] = w
 * (padding slots: env_mat = -mean / stddev, diff = 0, sw = 0).  diff and sw
 * may be NULL.  A neighbor index >= nall or a type outside [0, ntypes) gives
 * NaN in that row.  rcut_smth >= rcut is an error, as in the reference.
 * ------------------------------------------------------------------------ */
int dpmi_env_mat_pt_f64(double* env_mat, double* diff, double* sw, const double* ext_coord, const int* nlist,
                        const int* atype, const double* mean, const double* stddev, int64_t nframes, int64_t nloc,
                        int64_t nall, int64_t nnei, int ntypes, int ncomp, double rcut, double rcut_smth,
                        double protection, int use_exp_switch, void* stream);
int dpmi_env_mat_pt_f32(float* env_mat, float* diff, float* sw, const float* ext_coord, const int* nlist,
                        const int* atype, const float* mean, const float* stddev, int64_t nframes, int64_t nloc,
                        int64_t nall, int64_t nnei, int ntypes, int ncomp, double rcut, double rcut_smth,
                        double protection, int use_exp_switch, void* stream);
/* Coordinate gradient (the backward of the function above):
 * grad_coord [nframes, nall, 3] = sum over slots of the derivative of
 * <g_env_mat, env_mat> + <g_diff, diff> + <g_sw, sw> w.r.t. the extended
 * coordinates.  Any of g_env_mat / g_diff / g_sw may be NULL (zero).
 * grad_coord is overwritten.  Neighbor terms are added with atomics, so the
 * last bits depend on timing, as the reference's autograd scatter does. */
int dpmi_env_mat_pt_grad_f64(double* grad_coord, const double* g_env_mat, const double* g_diff, const double* g_sw,
                             const double* ext_coord, const int* nlist, const int* atype, const double* stddev,
                             int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,
                             double rcut, double rcut_smth, double protection, int use_exp_switch, void* stream);
int dpmi_env_mat_pt_grad_f32(float* grad_coord, const float* g_env_mat, const float* g_diff, const float* g_sw,
                             const float* ext_coord, const int* nlist, const int* atype, const float* stddev,
                             int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,
                             double rcut, double rcut_smth, double protection, int use_exp_switch, void* stream);
/* Backward of the coordinate gradient w.r.t. (g_env_mat, g_diff, g_sw)
 * applied to v = dL/d grad_coord [nframes, nall, 3] (the double backward of
 * an energy + force loss): gg_env_mat [.., ncomp] = J_env v, gg_diff [.., 3],
 * gg_sw [..] (each may be NULL).  The second derivative w.r.t. the
 * coordinates is not formed, as the reference's registered op gradients do
 * not form it. */
int dpmi_env_mat_pt_grad_grad_f64(double* gg_env_mat, double* gg_diff, double* gg_sw, const double* v,
                                  const double* ext_coord, const int* nlist, const int* atype, const double* stddev,
                                  int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,
                                  double rcut, double rcut_smth, double protection, int use_exp_switch, void* stream);
int dpmi_env_mat_pt_grad_grad_f32(float* gg_env_mat, float* gg_diff, float* gg_sw, const float* v,
                                  const float* ext_coord, const int* nlist, const int* atype, const float* stddev,
                                  int64_t nframes, int64_t nloc, int64_t nall, int64_t nnei, int ntypes, int ncomp,
                                  double rcut, double rcut_smth, double protection, int use_exp_switch, void* stream);

/* format_nlist_cpu (source/lib/src/fmt_nlist.cc:145-177; the formatting step
 * of forward_lower with an external list, make_model.py:334-415, and of the
 * LAMMPS / C++ path, DeepPotPT.cc:180-330): row i of cand (width ncand,
 * extended indices; numneigh[i] entries, or NULL to scan the whole row
 * skipping -1) is formatted like the search output: neighbors of type in
 * [0, ntypes) with float r^2 <= rc^2, by type section, sorted by (float r^2,
 * index), -1 padding, overflow as above.  Candidates must be distinct.
 * workspace: 4 bytes of device memory (a status word).  Waits for the stream
 * to read that status back. */
int dpmi_format_nlist_f64(int* nlist, int* overflow, int* nlist_mapped, const int* cand, const int* numneigh,
                          int64_t ncand, const double* ext_coord, const int* ext_type, const int* mapping,
                          int64_t nloc, int64_t nall, float rcut, const int* sel_host, int ntypes, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Force and virial.  prod_force_a_gpu / prod_virial_a_gpu
 * (source/lib/include/prod_force.h, prod_virial.h; CPU semantics
 * source/lib/src/prod_force.cc:23-76, prod_virial.cc:22-70).
 * force [nframes, nall, 3]; virial [nframes, 9]; atom_virial
 * [nframes, nall, 9] (may be NULL).  The virial reduction is deterministic
 * (per-block partials, then one compensated sum); its workspace size comes
 * from dpmi_prod_virial_workspace_bytes.
 * ------------------------------------------------------------------------ */
int dpmi_prod_force_se_a_f64(double* force, const double* net_deriv, const double* in_deriv,
                             const int* nlist, int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes,
                             void* stream);
int dpmi_prod_force_se_a_f32(float* force, const float* net_deriv, const float* in_deriv,
                             const int* nlist, int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes,
                             void* stream);
int dpmi_prod_virial_workspace_bytes(int64_t* bytes, int64_t nloc, int64_t nframes);
int dpmi_prod_virial_se_a_f64(double* virial, double* atom_virial, const double* net_deriv,
                              const double* in_deriv, const double* rij, const int* nlist, int64_t nloc,
                              int64_t nall, int64_t nnei, int64_t nframes, void* workspace,
                              int64_t workspace_bytes, void* stream);
int dpmi_prod_virial_se_a_f32(float* virial, float* atom_virial, const float* net_deriv,
                              const float* in_deriv, const float* rij, const int* nlist, int64_t nloc,
                              int64_t nall, int64_t nnei, int64_t nframes, void* workspace,
                              int64_t workspace_bytes, void* stream);

/* dpmi_prod_virial_se_a_f64 without a workspace argument: it allocates
 * nothing, so it suits callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Its per-workgroup partials
 * live in 8 MiB of device memory the library owns on each GPU (held until
 * the process ends): inputs need nframes * ceil(nloc / 4) <= 116,508, larger
 * ones return DPMI_ERR_ARG.  Every call on one GPU shares that memory, so
 * calls that overlap on one GPU (other streams, other threads) race; calls
 * queued on one stream do not.  Results are those of the workspace form. */
int dpmi_prod_virial_se_a_noalloc_f64(double* virial, double* atom_virial, const double* net_deriv,
                                      const double* in_deriv, const double* rij, const int* nlist, int64_t nloc,
                                      int64_t nall, int64_t nnei, int64_t nframes, void* stream);

/* ------------------------------------------------------------------------
 * Gradients of force and virial w.r.t. net_deriv (the training step's
 * double backward).  prod_force_grad_a_cpu (source/lib/src/
 * prod_force_grad.cc:22-79; TF op ProdForceSeAGrad,
 * source/op/tf/prod_force_grad_multi_device.cc) and prod_virial_grad_a_cpu
 * (source/lib/src/prod_virial_grad.cc:20-62; ProdVirialSeAGrad).
 * grad [nframes, nall, 3] (force) or [nframes, 9] (virial) -> grad_net
 * [nframes, nloc, 4 nnei]; in_deriv [nframes, nloc, 4 nnei, 3]; rij
 * [nframes, nloc, nnei, 3]; nlist [nframes, nloc, nnei].  The force
 * gradient has nall rows per frame and a neighbor index j >= nall reads
 * row j % nloc: with nall = nloc this is the reference op exactly (its
 * grad is [nframes, nloc * 3]); with nall = the extended count it is the
 * exact transpose of dpmi_prod_force_se_a.  As the reference's registered
 * gradients, the virial gradient takes the total virial's gradient only.
 * ------------------------------------------------------------------------ */
int dpmi_prod_force_grad_a_f64(double* grad_net, const double* grad, const double* in_deriv, const int* nlist,
                               int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes, void* stream);
int dpmi_prod_force_grad_a_f32(float* grad_net, const float* grad, const float* in_deriv, const int* nlist,
                               int64_t nloc, int64_t nall, int64_t nnei, int64_t nframes, void* stream);
int dpmi_prod_virial_grad_a_f64(double* grad_net, const double* grad, const double* in_deriv, const double* rij,
                                const int* nlist, int64_t nloc, int64_t nnei, int64_t nframes, void* stream);
int dpmi_prod_virial_grad_a_f32(float* grad_net, const float* grad, const float* in_deriv, const float* rij,
                                const int* nlist, int64_t nloc, int64_t nnei, int64_t nframes, void* stream);

/* Backward of the env-mat for the model path (the reverse-mode derivative
 * that deepmd/pt obtains from autograd through prod_env_mat,
 * deepmd/pt/model/descriptor/env_mat.py:11-92, followed by the ghost ->
 * local scatter of communicate_extended_output,
 * deepmd/pt/model/model/transform_output.py:317-383, and the virial of
 * transform_output.py:85-87).  net_deriv = dE/d em [nloc, nnei, 4]; the env-mat
 * derivative is recomputed from coordinates, never stored.  With mapping
 * non-NULL, force is [nloc, 3] (ghost contributions folded onto owners);
 * with mapping NULL, force is [nall, 3] on extended atoms.  force holds
 * -dE/dx (the sign of a force).  virial [9] (may be NULL) is the
 * deterministic compensated sum over pairs of g (x) r_ij.  Coordinates are
 * double; the _f32 variant evaluates the env-mat derivative in float.
 * Workspace: dpmi_prod_virial_workspace_bytes(nloc, 1). */
int dpmi_force_virial_recompute_f64(double* force, double* virial, const double* net_deriv, const int* nlist,
                                    const int* mapping, const double* ext_coord, const int* ext_type, int64_t nloc,
                                    int64_t nall, const double* std, float rcut, float rcut_smth,
                                    const int* sel_host, int ntypes, void* workspace, int64_t workspace_bytes,
                                    void* stream);
int dpmi_force_virial_recompute_f32(double* force, double* virial, const float* net_deriv, const int* nlist,
                                    const int* mapping, const double* ext_coord, const int* ext_type, int64_t nloc,
                                    int64_t nall, const float* std, float rcut, float rcut_smth,
                                    const int* sel_host, int ntypes, void* workspace, int64_t workspace_bytes,
                                    void* stream);

/* Cell-wise form of dpmi_force_virial_recompute_* for a list built with
 * cand_slot (same results up to summation order; deterministic).  Each
 * cell's neighbor terms accumulate in on-chip memory at the neighbor's slot
 * of the cell's candidate table and every local atom then sums its per-cell
 * partials, so the neighbor forces need no memory atomics (atoms the search
 * handled one by one fall back to atomics).  Replaces the same reference
 * functions as dpmi_force_virial_recompute_*; box_host and rcut as given to
 * dpmi_copy_coord_f64, nlist_workspace = that call's workspace.  Forces are
 * folded onto local atoms (mapping may be NULL only without ghosts). */
int dpmi_force_virial_cells_workspace_bytes(int64_t* bytes, const double* box_host, int64_t nloc, float rcut,
                                            int fp64);
int dpmi_force_virial_cells_f64(double* force, double* virial, const double* net_deriv, const int* nlist,
                                const int16_t* cand_slot, const int* mapping, const double* ext_coord,
                                const int* ext_type, int64_t nloc, int64_t nall, const double* std, float rcut,
                                float rcut_smth, const int* sel_host, int ntypes, const double* box_host,
                                const void* nlist_workspace, void* workspace, int64_t workspace_bytes,
                                void* stream);
int dpmi_force_virial_cells_f32(double* force, double* virial, const float* net_deriv, const int* nlist,
                                const int16_t* cand_slot, const int* mapping, const double* ext_coord,
                                const int* ext_type, int64_t nloc, int64_t nall, const float* std, float rcut,
                                float rcut_smth, const int* sel_host, int ntypes, const double* box_host,
                                const void* nlist_workspace, void* workspace, int64_t workspace_bytes,
                                void* stream);

/* ------------------------------------------------------------------------
 * Tabulated embedding fused with G^T R.  tabulate_fusion_se_a_gpu and its
 * _grad/_grad_grad (source/lib/include/tabulate.h:6-46; CPU semantics
 * source/lib/src/tabulate.cc:80-383).  table [nspline, 6*M] (coefficients
 * of channel k at 6k..6k+5), table_info_host: 6 values on the HOST as in
 * the reference op (source/op/pt/tabulate_multi_device.cc), em_x [nloc*nnei],
 * em [nloc, nnei, 4], out/dy/dz_dy [nloc, 4, M].  is_sorted enables the
 * reference's early exit at the first padded neighbor.
 * ------------------------------------------------------------------------ */
int dpmi_tabulate_fusion_se_a_f64(double* out, const double* table, const double* table_info_host,
                                  const double* em_x, const double* em, int64_t nloc, int64_t nnei,
                                  int64_t last_layer_size, int is_sorted, void* stream);
int dpmi_tabulate_fusion_se_a_f32(float* out, const float* table, const float* table_info_host,
                                  const float* em_x, const float* em, int64_t nloc, int64_t nnei,
                                  int64_t last_layer_size, int is_sorted, void* stream);
int dpmi_tabulate_fusion_se_a_grad_f64(double* dy_dem_x, double* dy_dem, const double* table,
                                       const double* table_info_host, const double* em_x, const double* em,
                                       const double* dy, int64_t nloc, int64_t nnei, int64_t last_layer_size,
                                       int is_sorted, void* stream);
int dpmi_tabulate_fusion_se_a_grad_f32(float* dy_dem_x, float* dy_dem, const float* table,
                                       const float* table_info_host, const float* em_x, const float* em,
                                       const float* dy, int64_t nloc, int64_t nnei, int64_t last_layer_size,
                                       int is_sorted, void* stream);
int dpmi_tabulate_fusion_se_a_grad_grad_f64(double* dz_dy, const double* table, const double* table_info_host,
                                            const double* em_x, const double* em, const double* dz_dy_dem_x,
                                            const double* dz_dy_dem, int64_t nloc, int64_t nnei,
                                            int64_t last_layer_size, int is_sorted, void* stream);
int dpmi_tabulate_fusion_se_a_grad_grad_f32(float* dz_dy, const float* table, const float* table_info_host,
                                            const float* em_x, const float* em, const float* dz_dy_dem_x,
                                            const float* dz_dy_dem, int64_t nloc, int64_t nnei,
                                            int64_t last_layer_size, int is_sorted, void* stream);

/* Table builder of the compressed embedding (DPTabulate.build,
 * deepmd/utils/tabulate.py:67-147, 273-367, with the value / slope /
 * curvature chain of deepmd/utils/tabulate_math.py:136-251, 353-470), FP64:
 * the embedding net of one neighbor type evaluated at the grid nodes xx
 * [nnodes] with its first and second derivative (node_vdd [3, nnodes, M],
 * written), then the quintic Hermite coefficients of every interval of width
 * tt [nnodes - 1]: table [nnodes - 1, 6 M] in the layout
 * dpmi_tabulate_fusion_se_a_f64 reads.  params: per layer w [nin][nout]
 * (row-major), b [nout], then the factor applied to tanh [nout] (the
 * timestep idt, or ones: the reference builder ignores idt); dims_host
 * [nlayers][2] = (nin, nout) with nin of the first layer 1 and widths
 * <= 1024, nlayers <= 8; nout == nin adds the layer input, nout == 2 nin adds
 * it twice, as MLPLayer does. */
int dpmi_tabulate_build_f64(double* table, double* node_vdd, const double* xx, const double* tt, int64_t nnodes,
                            const double* params, const int* dims_host, int nlayers, void* stream);

/* ------------------------------------------------------------------------
 * Fused embedding net + G^T R and the descriptor, bf16 (or fp16: the _f16
 * entry points) MFMA with fp32 accumulation (the precision form of
 * DescrptBlockSeA.forward,
 * deepmd/pt/model/descriptor/se_a.py:802-858, embedding layers of
 * deepmd/pt/model/network/mlp.py:188-219; type_one_side, widths
 * (n1, 2 n1, 4 n1), n1 <= 32, resnet with timestep).  The reference has no
 * single function for this: it chains torch MLP layers and matmuls.
 * frags: [ntypes, 20, 64, 8] bf16 (fp16 for _f16) weight fragments and consts
 * [ntypes, 576] fp32, both laid out by dpmi.embed.pack_embedding (internal
 * feature order q3; see that module).  em [nloc, nnei, 4] fp32.
 * ------------------------------------------------------------------------ */
/* counts [nloc, ntypes] int16 = listed neighbors of each atom per type
 * section of nlist [nloc, sum(sel)] (entries before the section's first -1,
 * where env_mat_a_cpu stops, source/lib/src/env_mat.cc:158-160). */
int dpmi_se_a_section_counts(int16_t* counts, const int* nlist, int64_t nloc, const int* sel_host, int ntypes,
                             void* stream);
/* gr_part [ntypes, nloc, 4 n1, 4]: sum over each type's neighbors of
 * (em^T G)^T (original feature order, the four components of a feature
 * contiguous).  Optional padding shortcut (counts non-NULL, as
 * dpmi_se_a_section_counts writes them): the slots after an atom's listed
 * neighbors must all hold the same em row, pad_em [ntypes (center),
 * ntypes (neighbor), 4] for the atom's type atype [nloc] (true when davg /
 * dstd are constant over each section, the reference's trailing-padding
 * assumption, source/lib/src/gpu/tabulate.cu:217-237); their share is then added as
 * n_pad em_pad^T G(em_pad) instead of evaluating the net per slot. */
int dpmi_se_a_embed_fwd_bf16(float* gr_part, const float* em, const void* frags, const float* consts, int64_t nloc,
                             const int* sel_host, int ntypes, int n1, const int16_t* counts, const int* atype,
                             const float* pad_em, void* stream);
int dpmi_se_a_embed_fwd_f16(float* gr_part, const float* em, const void* frags, const float* consts, int64_t nloc,
                            const int* sel_host, int ntypes, int n1, const int16_t* counts, const int* atype,
                            const float* pad_em, void* stream);
/* gr [nloc, 4, 4 n1] = sum_t gr_part / nnei; desc [nloc, 4 n1 * axis] =
 * GR^T GR[:, :axis] (se_a.py:851-858) in fp32 (desc_kind 0), bf16 (1) or
 * fp16 (2). */
int dpmi_se_a_descriptor_fwd(float* gr, void* desc, const float* gr_part, int64_t nloc, int ntypes, int nnei,
                             int n1, int axis, int desc_kind, void* stream);
/* dgr [nloc, 128, 4] = dE/d(sum_t gr_part) from ddesc = dE/d desc, in the
 * embedding kernels' internal feature order, the four components of a
 * feature contiguous (gr as written above). */
int dpmi_se_a_descriptor_bwd(float* dgr, const float* ddesc, const float* gr, int64_t nloc, int nnei, int n1,
                             int axis, void* stream);
/* net_deriv [nloc, nnei, 4] = dE/d em, recomputing the embedding forward
 * with its tangent in s = em[..., 0] (forward mode: dE/ds = sum_c em[c]
 * sum_q dGR[c][q] dG[q]/ds; n1 = width of the first embedding layer, as
 * packed).  With counts (the
 * padding shortcut of the forward) the rows after each section's listed
 * neighbors are written 0: they are padding slots, which no force term
 * reads.  The fp16 form stages each atom's dgr scaled by a power of two
 * (its largest magnitude brought into [1, 2)) and unscales the result,
 * exactly, so its fp16 dgr operands stay normal numbers. */
int dpmi_se_a_embed_bwd_bf16(float* net_deriv, const float* em, const float* dgr, const void* frags,
                             const float* consts, int64_t nloc, const int* sel_host, int ntypes, int n1,
                             const int16_t* counts, void* stream);
int dpmi_se_a_embed_bwd_f16(float* net_deriv, const float* em, const float* dgr, const void* frags,
                            const float* consts, int64_t nloc, const int* sel_host, int ntypes, int n1,
                            const int16_t* counts, void* stream);

/* FP64 form of the same two kernels (v_mfma_f64_16x16x4f64, fp64
 * throughout).  Weights as packed by dpmi.embed.pack_embedding_f64: w2
 * [ntypes, 32, 64], w3 [ntypes, 64, 128] in the kernels' internal feature
 * order, consts [ntypes, 480].  gr_part [ntypes, nloc, 4, 4 n1] (original
 * feature order); dgr [nloc, 4, 4 n1] = dE / d(sum over types of gr_part),
 * the same for every type; net_deriv [nloc, nnei, 4].  counts / atype /
 * pad_em (fp64): the padding shortcut of the bf16 kernels above. */
int dpmi_se_a_embed_fwd_f64(double* gr_part, const double* em, const double* w2, const double* w3,
                            const double* consts, int64_t nloc, const int* sel_host, int ntypes, int n1,
                            const int16_t* counts, const int* atype, const double* pad_em, void* stream);
int dpmi_se_a_embed_bwd_f64(double* net_deriv, const double* em, const double* dgr, const double* w2,
                            const double* w3, const double* consts, int64_t nloc, const int* sel_host, int ntypes,
                            int n1, const int16_t* counts, void* stream);

/* ------------------------------------------------------------------------
 * Fused energy fitting net (bf16 MFMA operands, or fp16 for the _f16
 * entry points, fp32 accumulation).
 * Replaces GeneralFitting._forward_common + its autograd backward
 * (deepmd/pt/model/task/fitting.py:734-900; FittingNet layers of
 * deepmd/dpmodel/utils/network.py:1031-1090): per type, three hidden layers
 * of width H <= 256 (the second and third with the skip connection) and a
 * linear H -> 1 output.  desc [nloc, k1] is the descriptor in bf16 (fp16)
 * as dpmi_se_a_descriptor_fwd writes it with desc_kind 1 (2).  frags /
 * consts / ebias are laid out by
 * dpmi.fitting.pack_fitting; perm [nloc] lists the local atoms ordered by
 * type and type_start [ntypes + 1] (device) delimits each type's run.
 * ------------------------------------------------------------------------ */
/* atom_energy [nloc] fp64 = network output + bias_atom_e + out_bias;
 * ddesc [nloc, k1] fp32 = dE/d desc for E = sum atom_energy. */
int dpmi_se_a_fitting_bf16(double* atom_energy, float* ddesc, const void* desc, const int* perm,
                           const int* type_start, const void* frags, const float* consts, const double* ebias,
                           int64_t nloc, int ntypes, int k1, void* stream);
int dpmi_se_a_fitting_f16(double* atom_energy, float* ddesc, const void* desc, const int* perm,
                          const int* type_start, const void* frags, const float* consts, const double* ebias,
                          int64_t nloc, int ntypes, int k1, void* stream);
/* The same fitting net with the descriptor fused in (axis 16, M = 4 n1 with
 * 16 <= M <= 128): gr_part [ntypes, nloc, M, 4] as dpmi_se_a_embed_fwd_bf16
 * writes it; each 32-atom workgroup forms GR = sum_t gr_part / nnei and
 * D = GR^T GR[:, :16] (se_a.py:851-858) in LDS and, instead of dE/dD, writes
 * dgr [nloc, 128, 4] = dE/d(sum_t gr_part) in the embedding kernels' internal
 * feature order (what dpmi_se_a_descriptor_bwd produces, the input of
 * dpmi_se_a_embed_bwd_bf16).  Replaces the descriptor kernels and
 * dpmi_se_a_fitting_bf16 on that path; frags / consts / ebias as above with
 * k1 = 16 M.  GR is held in LDS as fp32 for M <= 104 and rounded to fp16
 * there for 104 < M <= 128 (the D products and sums stay fp32). */
int dpmi_se_a_fitting_gr_bf16(double* atom_energy, float* dgr, const float* gr_part, const int* perm,
                              const int* type_start, const void* frags, const float* consts, const double* ebias,
                              int64_t nloc, int ntypes, int n1, int nnei, void* stream);
int dpmi_se_a_fitting_gr_f16(double* atom_energy, float* dgr, const float* gr_part, const int* perm,
                             const int* type_start, const void* frags, const float* consts, const double* ebias,
                             int64_t nloc, int ntypes, int n1, int nnei, void* stream);
/* Number of 8 x 16-bit lanes (64 per fragment) of one type's packed fitting net. */
int dpmi_se_a_fitting_frag_count(int64_t* count, int k1);

/* FP64 form of the fused descriptor + fitting kernel (v_mfma_f64_16x16x4f64,
 * FP64 throughout): replaces the descriptor of DescrptBlockSeA.forward
 * (deepmd/pt/model/descriptor/se_a.py:851-858), GeneralFitting._forward_common
 * (deepmd/pt/model/task/fitting.py:734-900) and its autograd backward.
 * gr_part [nparts, nloc, 4, M] (GR = sum over the parts / nnei; the parts are
 * the per-neighbor-type G^T R of the embedding or of tabulate_fusion_se_a);
 * D = GR^T GR[:, :16] (axis 16, 16 <= M <= 128) is formed inside the kernel,
 * never stored.  Hidden width H <= 256, three hidden layers as the bf16
 * kernel.  perm / type_start as above.  frags: per type
 * dpmi_se_a_fitting_f64_frag_doubles doubles, consts [ntypes, 7*256 + 1],
 * ebias [ntypes], laid out by dpmi.fitting.pack_fitting_f64.
 * atom_energy [nloc]; dgr [nloc, 4, M] = dE/d(sum of the parts) for
 * E = sum atom_energy (the input of dpmi_se_a_embed_bwd_f64 and the dy of
 * dpmi_tabulate_fusion_se_a_grad_f64).  Deterministic (fixed-order sums). */
int dpmi_se_a_fitting_gr_f64(double* atom_energy, double* dgr, const double* gr_part, int nparts, const int* perm,
                             const int* type_start, const double* frags, const double* consts, const double* ebias,
                             int64_t nloc, int ntypes, int M, int H, int nnei, void* stream);
int dpmi_se_a_fitting_f64_frag_doubles(int64_t* count, int M, int H);

/* energy [1] = sum of atom_energy [n], FP64 with compensated (Neumaier)
 * summation in a fixed order (the per-frame energy reduction of
 * deepmd/pt/model/model/transform_output.py:165,181). */
int dpmi_energy_sum_f64(double* energy, const double* atom_energy, int64_t n, void* stream);

/* ------------------------------------------------------ training step
 * The optimizer side of the training step (deepmd/pt/train/training.py:
 * 1393-1460), with the step count and the rate kept in device memory so
 * that the whole step can be captured and replayed as a HIP graph.  Neither
 * function waits on the stream. */

/* Learning-rate tick at the start of a step: with k = *step (steps begun so
 * far, int64 on the device) writes hyper [4] (device) = [lr_k, pref_e,
 * pref_f, k + 1] and sets *step = k + 1.  lr_k follows the LambdaLR over
 * LearningRateExp that the reference trainer builds (training.py:1138-1150;
 * deepmd/dpmodel/utils/learning_rate.py:167-217, 367-397); the prefactors
 * follow lr as EnergyStdLoss does (deepmd/pt/loss/ener.py:229-231).
 * sched_host (host) [11] = start_lr, decay_rate, stop_lr, decay_steps,
 * warmup_steps, warmup_start_lr, smooth (0/1), start_pref_e, limit_pref_e,
 * start_pref_f, limit_pref_f; its values are read at the call. */
int dpmi_lr_tick(double* hyper, int64_t* step, const double* sched_host, void* stream);

/* One Adam step over n elements of a flat parameter buffer, in place on
 * param, exp_avg and exp_avg_sq, with lr = hyper[0] and t = hyper[3] read on
 * the device: replaces torch.optim.Adam(fused=True) of the reference trainer
 * (training.py:1028-1030, stepped at 1451-1452; no amsgrad, no maximize).
 * exp_avg / exp_avg_sq start at zero. */
int dpmi_adam_f64(double* param, const double* grad, double* exp_avg, double* exp_avg_sq, int64_t n,
                  const double* hyper, double beta1, double beta2, double eps, double weight_decay, void* stream);
int dpmi_adam_f32(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                  const double* hyper, double beta1, double beta2, double eps, double weight_decay, void* stream);

/* out [F] = column sums of x [R, F] (row-major) in a fixed order (two
 * passes; part: workspace of 256 * F elements).  The gradient of a bias or
 * timestep broadcast over the R neighbor rows of an embedding layer
 * (MLPLayer, deepmd/pt/model/network/mlp.py:188-219) in the training step,
 * deterministic and free of memory-set nodes when captured in a graph; the
 * reference leaves this reduction to torch autograd. */
int dpmi_row_sum_f64(double* out, const double* x, double* part, int64_t R, int64_t F, void* stream);
int dpmi_row_sum_f32(float* out, const float* x, float* part, int64_t R, int64_t F, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DPMI_H_ */
