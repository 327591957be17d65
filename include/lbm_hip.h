/*
 * lbm_hip.h - C ABI of the MI355X (gfx950) lattice Boltzmann library liblbm_hip.so.
 *
 * Every entry point takes plain device pointers (allocated by the caller),
 * sizes and scalars, launches on the given HIP stream (NULL = default stream)
 * without synchronising, and returns LBM_OK or a negative LBM_ERR_* code;
 * lbm_last_error() then describes the failure. The library never frees
 * caller memory.
 *
 * Populations are stored structure-of-arrays, f[q][z][y][x] (x fastest,
 * 2D grids have nz = 1), shifted by the lattice weights (f~ = f - w), in
 * float (LBM_F32) or double (LBM_F64). kind[] holds one byte per node:
 * bit 0 solid, bit 1 IBM-forced fluid node, bit 2 fluid node with a solid
 * neighbour at some x - c_i (set by the host, lets interior wavefronts skip
 * the bounce-back checks). wall_u and force fields are [d][z][y][x] arrays
 * of the population precision; wall_u (and the batched wall pool) may be NULL
 * when every wall is at rest: the moving-wall terms are then exact zeros and
 * no wall velocity is read.
 *
 * Each declaration names the reference interface it replaces
 * (ShauryaAttreya37/High-Performance-LBM-On-GPU-, paths relative to its root).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_API __attribute__((visibility("default")))

enum { LBM_D2Q9 = 0, LBM_D3Q19 = 1 };
enum { LBM_F32 = 0, LBM_F64 = 1 };
/* Force handling: uniform F0; F0 plus a sparse IBM field at forced nodes;
 * F0 plus a dense per-node body force (plus the IBM field at forced nodes). */
enum { LBM_FORCE_UNIFORM = 0, LBM_FORCE_FIELD = 1, LBM_FORCE_DENSE = 2 };
enum {
    LBM_OK = 0,
    LBM_ERR_INVALID = -1,  /* bad argument */
    LBM_ERR_HIP = -2,      /* HIP runtime error */
    LBM_ERR_INTERNAL = -3
};

/* Message of the last failed call on this thread ("" if none). */
LBM_API const char* lbm_last_error(void);
/* Library version string. No counterpart in the reference. */
LBM_API const char* lbm_version(void);

/* Host-side copy of the compiled lattice tables: c[3*Q] (x, y, z rows), opp[Q], w[Q].
 * Replaces lattice.cuda_preamble (src/lbm/lattice.py:174-214); needs no GPU. */
LBM_API int lbm_lattice_tables(int lattice, int* q, int* c, int* opp, double* w);

/* Same tables read back from device code (one tiny kernel) into device buffers
 * c[3*Q] (int32), opp[Q] (int32), w[Q] (dtype).
 * Replaces the on-device constant check of tests/test_cuda_backend.py:160-183. */
LBM_API int lbm_lattice_tables_device(int lattice, int dtype, int* c, int* opp, void* w, void* stream);

/* Stream-collide entry points (lbm_pull_step, lbm_aa_even, lbm_aa_odd, lbm_lds_step):
 * groups_per_cu > 0 caps the resident workgroups per compute unit (requested as unused
 * LDS: the bandwidth of the AA pair at 256^3 peaks with 2 groups of 256 x 2 per CU);
 * 0 leaves occupancy to the hardware. Blocks have at most 1024 threads, at most 512 for
 * float64 D3Q19 (those kernels are compiled for 512 lanes to keep all populations in
 * registers); larger blocks are rejected with LBM_ERR_INVALID before any launch. */

/* Fused pull stream + halfway bounce-back (Ladd) + TRT/Guo collision, two lattices.
 * Replaces stream_collide_naive (src/lbm/backends/kernels/stream_collide_naive.cu:15-49),
 * launched from CudaSolver._advance (src/lbm/backends/cuda_solver.py:279-284).
 * mask_mode 0: no solid/forced nodes anywhere (kind ignored, may be NULL); 1: use kind.
 * force_field: IBM field (FIELD/DENSE modes, read at forced nodes); body_field:
 * dense per-node force (DENSE mode). bx * by <= 1024 threads per block. */
LBM_API int lbm_pull_step(int lattice, int dtype, const void* f_src, void* f_dst,
                          const uint8_t* kind, const void* wall_u, const void* force_field,
                          const void* body_field, int nx, int ny, int nz, double omega_plus,
                          double omega_minus, double fx, double fy, double fz, int force_mode,
                          int mask_mode, int bx, int by, int groups_per_cu, void* stream);

/* lbm_pull_step (mask_mode 1) with the link rules evaluated inside the step and, optionally,
 * the momentum-exchange force of the state it reads, in one launch.
 * Link rules: rule rows in the pull layout (as for lbm_rules_apply: term-major slot/coeff,
 * row stride n_rules, m = rule_m terms, constant) sorted by (node, direction j), where the
 * row of link (node, j) gives f_j(node, t + 1). The rule_nodes nodes with rows, sorted, are
 * described by rule_node_info[4 k .. 4 k + 3] = {node, bit set of the directions i whose
 * upstream node node - c_i is solid, bit set of the directions j with a row, index of the
 * node's first row} (16-byte aligned), with rule_node_first[k] = that first row again
 * (rule_node_first[rule_nodes] = n_rules). kind must carry KIND_RULES (8) on exactly the
 * listed nodes: those are updated, 16 to 64 nodes per workgroup, by an extra layer of
 * workgroups that evaluate the nodes' rows (one row per thread, the values shared through
 * LDS; bx * by >= 64), every other non-solid node by the ordinary grid. f_src is never
 * written during the step, so the rows read the populations a pre-pass reads and the
 * result equals lbm_rules_apply + lbm_pull_step bit for bit, without the pre-pass launch.
 * n_rules = 0 (rule_nodes = 0): no rules (the rule arrays may be NULL).
 * Force: when n_force > 0, the extra layer also reduces the force rows on f_src exactly as
 * lbm_affine_force does (same groups, same sums; bx * by must be a multiple of 64)
 * into force_out[d] (up to 2048 rows) or force_partial (3 doubles per group of
 * lbm_affine_force_groups; lbm_force_totals adds them).
 * Replaces _RulePass.__call__ followed by stream_collide_naive
 * (src/lbm/backends/cuda_solver.py:157-176, 279-284, stream_collide_naive.cu:15-49) and,
 * for per-step drag, ForceProbe.evaluate (src/lbm/forces.py:98-103). */
LBM_API int lbm_pull_step_links(int lattice, int dtype, const void* f_src, void* f_dst, const uint8_t* kind,
                                const void* wall_u, const void* force_field, const void* body_field, int nx, int ny,
                                int nz, double omega_plus, double omega_minus, double fx, double fy, double fz,
                                int force_mode, const uint32_t* rule_node_info, const int32_t* rule_node_first,
                                int rule_nodes, const int64_t* rule_slot,
                                const void* rule_coeff, const void* rule_constant, int64_t n_rules, int rule_m,
                                const int64_t* force_slot, const void* force_coeff, const void* force_constant,
                                const double* force_weight, int64_t n_force, int force_m, double* force_out,
                                double* force_partial, int bx, int by, int groups_per_cu, void* stream);

/* AA pattern, even half-step: read F[i](x), collide, write F[opp i](x).
 * Replaces aa_even (src/lbm/backends/kernels/stream_collide_aa.cu:33-57). */
LBM_API int lbm_aa_even(int lattice, int dtype, void* f, const uint8_t* kind,
                        const void* force_field, const void* body_field, int nx, int ny, int nz,
                        double omega_plus, double omega_minus, double fx, double fy, double fz,
                        int force_mode, int mask_mode, int bx, int by, int groups_per_cu, void* stream);

/* AA pattern, odd half-step: pull F[opp i](x - c_i), collide, push to F[i](x + c_i);
 * solid targets are reflected into F[opp i](x) with the wall term.
 * Replaces aa_odd (src/lbm/backends/kernels/stream_collide_aa.cu:59-98). */
LBM_API int lbm_aa_odd(int lattice, int dtype, void* f, const uint8_t* kind, const void* wall_u,
                       const void* force_field, const void* body_field, int nx, int ny, int nz,
                       double omega_plus, double omega_minus, double fx, double fy, double fz,
                       int force_mode, int mask_mode, int bx, int by, int groups_per_cu, void* stream);

/* Two-lattice step with each block's populations plus a one-node halo staged in LDS
 * (2D only; tile = block). Replaces stream_collide_tiled
 * (src/lbm/backends/kernels/stream_collide_tiled.cu:24-79). */
LBM_API int lbm_lds_step(int lattice, int dtype, const void* f_src, void* f_dst,
                         const uint8_t* kind, const void* wall_u, const void* force_field,
                         const void* body_field, int nx, int ny, int nz, double omega_plus,
                         double omega_minus, double fx, double fy, double fz, int force_mode,
                         int mask_mode, int bx, int by, int groups_per_cu, void* stream);

/* Shifted equilibrium from drho and u (device arrays of dtype, n nodes each; uz
 * ignored in 2D) into f; aa != 0 stores direction i in slot opp(i) (post-even layout).
 * Replaces Solver._initial_populations + CudaSolver.initialize
 * (src/lbm/solver.py:101-117, src/lbm/backends/cuda_solver.py:258-267). */
LBM_API int lbm_init_equilibrium(int lattice, int dtype, void* f, const void* drho, const void* ux,
                                 const void* uy, const void* uz, int64_t n, int aa, void* stream);

/* rho and u (rho[n], u[d][n], dtype) from stored post-collision populations; solid
 * nodes give rho = 1, u = 0. aa != 0 reads the post-even AA layout directly.
 * force_mode / force_field as for the step kernels (the field enters at forced nodes).
 * Replaces Solver.macroscopic (src/lbm/solver.py:209-221, diagnostics.py:72-80). */
LBM_API int lbm_macroscopic(int lattice, int dtype, const void* f, const uint8_t* kind,
                            const void* force_field, int64_t n, int aa, double fx, double fy,
                            double fz, int force_mode, void* rho, void* u, void* stream);

/* Divergence watchdog: out[0] = smallest flat index of a fluid node with a
 * non-finite drho or drho <= -1 (unchanged if none), out[1] |= 1 if any fluid
 * drho is non-finite. The caller initialises out = {UINT32_MAX, 0}.
 * Replaces Solver._check_health (src/lbm/solver.py:162-172). */
LBM_API int lbm_health(int lattice, int dtype, const void* f, const uint8_t* kind, int64_t n,
                       int aa, uint32_t* out, void* stream);

/* Link-rule pre-pass: value[r] = constant[r] + sum_t coeff[t][r] f[slot[t][r]] for all
 * rules first (pairwise sum over the m padded terms, as NumPy's row sum), then
 * f[target[r]] = value[r]. slot/coeff are term-major (m x n_rules); scratch holds
 * n_rules values of dtype (unused when n_rules <= 4096: one fused launch).
 * Replaces rules_evaluate / rules_write (src/lbm/backends/kernels/link_rules.cu:15-43)
 * and _RulePass.__call__ (src/lbm/backends/cuda_solver.py:157-176). */
LBM_API int lbm_rules_apply(int dtype, void* f, const int64_t* slot, const void* coeff,
                            const void* constant, const int64_t* target, int64_t n_rules, int m,
                            void* scratch, void* stream);

/* lbm_rules_apply without a scratch argument and without allocating: above 4096 rules the
 * rule values go through the library's own device buffer (one per GPU, 2^20 values = 8 MiB,
 * held until the process ends) instead of a caller buffer. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). At most 2^20 rules (LBM_ERR_INVALID above, before
 * any GPU work). Every call on one GPU shares that buffer: calls that overlap on one GPU
 * (other streams, other host threads) race. Same bits as lbm_rules_apply. */
LBM_API int lbm_rules_apply_noalloc(int dtype, void* f, const int64_t* slot, const void* coeff,
                                    const void* constant, const int64_t* target, int64_t n_rules, int m,
                                    void* stream);

/* lbm_rules_apply for a concatenation of rule sets (no counterpart in the reference): row r
 * sums only its first width[r] padded terms (its own set's width; m is the padded width of
 * the arrays), so every row rounds as when its set is applied alone. Always two launches
 * (evaluate into scratch[n_rules], then write). Used by the batched multi-lattice step. */
LBM_API int lbm_rules_apply_widths(int dtype, void* f, const int64_t* slot, const void* coeff,
                                   const void* constant, const int64_t* target, const int* width,
                                   int64_t n_rules, int m, void* scratch, void* stream);

/* The pre-pass of lbm_rules_apply in one launch, one rule per lane, no scratch, for a rule
 * set in which no rule reads a slot that another rule writes (a rule may read its own
 * target). The caller checks that property on the compiled set (the Python host does it
 * once per layout); under it every rule reads pre-rule populations in any lane order, so
 * the result equals lbm_rules_apply's bit for bit. width (may be NULL) gives row r its own
 * padded term count, as in lbm_rules_apply_widths. Same role as _RulePass.__call__
 * (src/lbm/backends/cuda_solver.py:157-176), which always runs its two kernels. */
LBM_API int lbm_rules_apply_direct(int dtype, void* f, const int64_t* slot, const void* coeff,
                                   const void* constant, const int* width, const int64_t* target,
                                   int64_t n_rules, int m, void* stream);

/* Values of compiled affine expressions on the state: out[r] (dtype), same
 * arithmetic as the rule evaluation. Replaces Solver.evaluate_affine
 * (src/lbm/solver.py:178-188) and link_rules.evaluate_affine (src/lbm/link_rules.py:269-272). */
LBM_API int lbm_affine_values(int dtype, const void* f, const int64_t* slot, const void* coeff,
                              const void* constant, int64_t n_rows, int m, void* out, void* stream);

/* Momentum-exchange force: out[a] (double) = sum_r weight[r][a] * value[r] over the
 * affine rows (value rounded to dtype, then widened), d components, reduced in a fixed
 * order (lane-strided partial sums, then a tree) that depends on n_rows only, so repeated
 * calls give the same bits. Workgroups have 1024 lanes; up to 2048 rows one writes out
 * directly; above, one workgroup per 1024 rows (at most 1024) writes partial sums into `scratch`
 * (lbm_affine_force_scratch_size bytes; may be NULL up to 2048 rows) and a second launch
 * adds them. Calls that share one scratch buffer must not overlap. Replaces
 * ForceProbe.evaluate's host matmul (src/lbm/forces.py:98-103) so a per-step drag
 * history stays on the device. */
LBM_API int lbm_affine_force(int dtype, const void* f, const int64_t* slot, const void* coeff,
                             const void* constant, const double* weight, int64_t n_rows, int m,
                             int d, double* out, void* scratch, void* stream);

/* The force of one step (lbm_affine_force with one workgroup: n_rows <= 2048) and the
 * link-rule pre-pass of the next step (lbm_rules_apply's one-workgroup form:
 * n_rules <= 4096) in one launch: both read the same state, every read precedes the rule
 * writes, and each gives the bits of its separate launch. No counterpart in the reference
 * (it evaluates ForceProbe on the host, src/lbm/forces.py:98-103, and runs _RulePass,
 * src/lbm/backends/cuda_solver.py:157-176, as its own launches); saves a launch per step
 * for per-step drag histories of small lattices. */
LBM_API int lbm_affine_force_rules(int dtype, void* f, const int64_t* force_slot, const void* force_coeff,
                                   const void* force_constant, const double* weight, int64_t n_rows,
                                   int force_m, int d, double* out, const int64_t* rule_slot,
                                   const void* rule_coeff, const void* rule_constant,
                                   const int64_t* rule_target, int64_t n_rules, int rule_m, void* stream);

/* Number of reduction groups of lbm_affine_force for n_rows rows (1 up to 2048 rows, then
 * one per 1024 rows up to 1024). No counterpart in the reference. Host only. */
LBM_API int lbm_affine_force_groups(int64_t n_rows, int* groups);

/* Totals of a history of force partial sums: out[s][a] (a < d) from partial[s][groups][3],
 * s < rows, added in lbm_affine_force's order (its second pass), so each row equals what
 * lbm_affine_force returns for the same state. No counterpart in the reference (per-step
 * drag histories on the device). */
LBM_API int lbm_force_totals(const double* partial, int64_t rows, int groups, int d, double* out, void* stream);

/* Bytes of device scratch lbm_affine_force needs for n_rows rows (0 up to 2048). No
 * counterpart in the reference (its force sum is a host matmul). Host only. */
LBM_API int lbm_affine_force_scratch_size(int64_t n_rows, int64_t* bytes);

/* Bytes of device scratch lbm_ibm_force needs for n_nodes unique stencil nodes
 * and n_markers markers. No counterpart in the reference (its IBM runs as array code on
 * CuPy temporaries, src/lbm/ibm.py:159-214). Host only. */
LBM_API int lbm_ibm_scratch_size(int lattice, int dtype, int n_nodes, int n_markers, int64_t* bytes);

/* Immersed-boundary direct forcing, all multi-direct iterations in one launch
 * (one workgroup; a body whose intermediates do not fit its LDS takes the many-workgroup
 * launches of lbm_ibm_force_multi when no record is asked for). Moment rows ((d+1) * n_nodes affine rows, term-major m-padded,
 * compiled for the current layout) give drho and rho u at the unique stencil nodes
 * nodes[n_nodes]; inverse[n_markers][3^d] and weights[n_markers][3^d] describe the
 * Roma stencils, areas[n_markers] the surface elements; spread_start[n_nodes + 1] /
 * spread_entry[] list, per node, the flat (marker, stencil) entries touching it in
 * increasing order. Writes field[a][nodes[n]] (d x n_total, dtype) and
 * body_force[d] (dtype) = -(sum of marker forces over all iterations). record (optional,
 * lbm_ibm_record_size bytes) receives the per-iteration intermediates for the adjoint; the
 * outputs do not depend on it.
 * Replaces ImmersedForcing.force_field (src/lbm/ibm.py:189-214) and the field
 * scatter of CudaSolver._pre_pass (src/lbm/backends/cuda_solver.py:286-292). */
LBM_API int lbm_ibm_force(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                          const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                          const void* weights, const void* areas, int n_markers, const int* spread_start,
                          const int* spread_entry, int iterations, double fx, double fy, double fz,
                          int64_t n_total, void* field, void* body_force, void* scratch, void* record,
                          void* stream);

/* lbm_ibm_force (without record) in 2 + 2 * iterations launches of many workgroups: per
 * stencil node the moment rows and u*, per marker the interpolation and marker force, per
 * node the spread, each value by the same expression in the same order, so the outputs equal
 * lbm_ibm_force's bit for bit. lbm_ibm_force takes this path itself when the body's
 * intermediates do not fit one workgroup's LDS (3D spheres); scratch as for lbm_ibm_force.
 * No counterpart in the reference (see lbm_ibm_force). */
LBM_API int lbm_ibm_force_multi(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                                const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                                const void* weights, const void* areas, int n_markers, const int* spread_start,
                                const int* spread_entry, int iterations, double fx, double fy, double fz,
                                int64_t n_total, void* field, void* body_force, void* scratch, void* stream);

/* Whether lbm_ibm_pull_step takes this body: D2Q9 (float32 or float64), bx * by in
 * 64 .. 1024 lanes, at least one lane per stencil node, and the one-workgroup forcing
 * (with `record` slots when record = 1) in the LDS (*ok = 1). No counterpart in the
 * reference. Host only. */
LBM_API int lbm_ibm_pull_step_supported(int lattice, int dtype, int n_nodes, int n_markers, int iterations,
                                        int record, int bx, int by, int* ok);

/* One forward step of a body-forced lattice in one launch: lbm_ibm_force of the state src
 * (field, body_force and record as there; scratch not needed) followed by
 * lbm_pull_step_links from src to dst with that field (force mode 1 or 2, link rules as
 * there, no force rows), with the same outputs bit for bit. The stencil nodes' own pull
 * update runs in the forcing's workgroup after the forcing, every other node in the other
 * workgroups; `kind` must carry bit 16 on exactly the stencil nodes (which the other
 * workgroups then skip), and no stencil node may be solid, next to a solid node or a rule
 * node. rows_in_lanes = 1 declares that the moment rows are those of lbm.ibm.moment_rows in
 * the pull layout (row a n_nodes + k of stencil node k: density, every direction in lattice
 * order with coefficient 1; momentum a, the directions with c_i,a != 0 in lattice order with
 * coefficient c_i,a, zero-padded to q terms; constant 0): each stencil node's lane then sums
 * them from the populations it pulls instead of gathering them (the same bits).
 * Replaces one iteration of the reference's pull loop: its pre-pass (link rules, then
 * the IBM force field and its scatter; src/lbm/backends/cuda_solver.py:286-292,
 * src/lbm/ibm.py:189-214) and its kernel launch (src/lbm/backends/cuda_solver.py:279-284). */
LBM_API int lbm_ibm_pull_step(int lattice, int dtype, const void* src, void* dst, const uint8_t* kind,
                              const void* wall, void* field, const void* body, int nx, int ny, int nz,
                              double omega_plus, double omega_minus, double fx, double fy, double fz, int force_mode,
                              const int64_t* slot, const void* coeff, const void* constant, int m,
                              const int64_t* nodes, int n_nodes, const int* inverse, const void* weights,
                              const void* areas, int n_markers, const int* spread_start, const int* spread_entry,
                              int iterations, void* body_force, void* record, const uint32_t* rule_node_info,
                              const int32_t* rule_node_first, int rule_nodes, const int64_t* rule_slot,
                              const void* rule_coeff, const void* rule_constant, int64_t n_rules, int rule_m,
                              int rows_in_lanes, int bx, int by, void* stream);

/* Bytes of one step's record of the forcing intermediates (optional `record` argument of
 * lbm_ibm_force, read back by lbm_ibm_adjoint), a multiple of 16; records at 16-byte aligned
 * addresses are copied with 16-byte accesses. No counterpart in the reference. Host only. */
LBM_API int lbm_ibm_record_size(int lattice, int dtype, int n_nodes, int n_markers, int iterations,
                                int64_t* bytes);

/* Copy of `bytes` (multiple of 16, 16-byte aligned buffers) with
 * 16-byte accesses: the device-copy bandwidth probe. Replaces the cp.copyto peak
 * measurement of benchmarks/bench_mlups.py:66-79. */
LBM_API int lbm_device_copy(void* dst, const void* src, int64_t bytes, void* stream);

/* Adjoint of one pull step (no counterpart in the reference, which has no adjoint):
 * lam_out = (d f^{t+1} / d f^t)^T lam_next for the collision and streaming part. f is the
 * state the forward pull step read (link-rule values already in place), force arguments as
 * for lbm_pull_step. The adjoint populations are pushed to x - c_i (bounce-back links to the
 * node's own opposite slot); lam_out must differ from lam_next and be zero at solid nodes.
 * grad_body[d][n] (double, optional) += dL/dF; nu_forced (dtype, optional) receives dL/dF at
 * IBM-forced nodes. bx * by <= 256. */
LBM_API int lbm_adj_pull_step(int lattice, int dtype, const void* lam_next, void* lam_out, const void* f,
                              const uint8_t* kind, const void* wall_u, const void* force_field,
                              const void* body_field, int nx, int ny, int nz, double omega_plus,
                              double omega_minus, double fx, double fy, double fz, int force_mode,
                              int mask_mode, double* grad_body, void* nu_forced, int bx, int by, void* stream);

/* lbm_adj_pull_step plus, in the same launch, an independent link-rule pre-pass on another
 * state buffer rule_f: as lbm_rules_apply in one workgroup (1 <= n_rules <= 4 * bx * by), or
 * with rules_direct = 1 as lbm_rules_apply_direct, one rule per lane of an extra grid layer
 * (n_rules up to the lanes of one z layer of the grid; only for sets in which no rule reads a
 * slot another rule writes). bx * by up to 1024 for D2Q9 float32, 512 for D2Q9 float64 and 256
 * for D3Q19. No counterpart in the reference: the checkpointed adjoint uses it to run the
 * forward recompute of one segment under the backward sweep of the next, saving a launch per
 * step. */
LBM_API int lbm_adj_pull_step_rules(int lattice, int dtype, const void* lam_next, void* lam_out, const void* f,
                                    const uint8_t* kind, const void* wall_u, const void* force_field,
                                    const void* body_field, int nx, int ny, int nz, double omega_plus,
                                    double omega_minus, double fx, double fy, double fz, int force_mode,
                                    int mask_mode, double* grad_body, void* nu_forced, int bx, int by,
                                    void* rule_f, const int64_t* rule_slot, const void* rule_coeff,
                                    const void* rule_constant, const int64_t* rule_target, int64_t n_rules,
                                    int rule_m, int rules_direct, void* stream);

/* Transposed link-rule pre-pass (no counterpart in the reference): a_r = lam[target_r],
 * lam[target_r] = 0, then lam[uslot[u]] += sum_{e in [ustart[u], ustart[u+1])} ucoeff[e] a[urule[e]]
 * over the distinct term slots. One workgroup; scratch holds n_rules values of dtype. */
LBM_API int lbm_adj_rules(int dtype, void* lam, const int64_t* target, int n_rules, const int64_t* uslot,
                          const int* ustart, const int* urule, const void* ucoeff, int n_unique,
                          void* scratch, void* stream);

/* Scratch bytes of lbm_ibm_adjoint. No counterpart in the reference. Host only. */
LBM_API int lbm_ibm_adjoint_scratch_size(int lattice, int dtype, int n_nodes, int n_markers, int iterations,
                                         int64_t* bytes);

/* Adjoint of lbm_ibm_force (no counterpart in the reference): takes the multi-direct
 * intermediates from record (written by lbm_ibm_force for the same f and markers) or, when
 * record is null, recomputes them from f, and runs the iterations backwards with seeds nu (dL/dh at forced nodes, d x n_total)
 * and body_weight[d] (dL/d body force, double). Adds the transposed moment rows into lam
 * (distinct slots tslot[n_tslots] with (trow, tcoeff) lists in tstart order) and accumulates
 * dL/dw into grad_w[n_markers][3^d] (double). When n_rules > 0 it first applies the
 * transposed link rules exactly as lbm_adj_rules does (same argument meaning), so one
 * launch covers both sparse adjoints of a step. slots_disjoint = 1 states that no rule term
 * slot (rule_uslot) is a rule target or a transposed moment slot (tslot): the kernel may then
 * read every lam value it updates before its first barrier; 0 is always correct. One
 * workgroup. */
LBM_API int lbm_ibm_adjoint(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                            const void* constant, int m, const int64_t* nodes, int n_nodes, const int* inverse,
                            const void* weights, const void* areas, int n_markers, const int* spread_start,
                            const int* spread_entry, int iterations, double fx, double fy, double fz,
                            int64_t n_total, const void* nu, const double* body_weight, void* lam,
                            const int64_t* tslot, const int* tstart, const int* trow, const void* tcoeff,
                            int n_tslots, double* grad_w, void* scratch, const int64_t* rule_target,
                            int n_rules, const int64_t* rule_uslot, const int* rule_ustart,
                            const int* rule_urule, const void* rule_ucoeff, int n_rule_unique,
                            void* rule_scratch, int slots_disjoint, const void* record, void* stream);

/* lbm_ibm_adjoint plus, in the same launch, an independent pull step (as lbm_pull_step, with
 * the uniform force fx, fy, fz of the IBM arguments and an n_total-node grid) from pull_src to
 * pull_dst, buffers other than f and lam. No counterpart in the reference: the checkpointed
 * adjoint uses it to run the forward recompute of one segment under the backward sweep of the
 * next, saving a launch per step. */
LBM_API int lbm_ibm_adjoint_pull(int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff,
                                 const void* constant, int m, const int64_t* nodes, int n_nodes,
                                 const int* inverse, const void* weights, const void* areas, int n_markers,
                                 const int* spread_start, const int* spread_entry, int iterations, double fx,
                                 double fy, double fz, int64_t n_total, const void* nu, const double* body_weight,
                                 void* lam, const int64_t* tslot, const int* tstart, const int* trow,
                                 const void* tcoeff, int n_tslots, double* grad_w, void* scratch,
                                 const int64_t* rule_target, int n_rules, const int64_t* rule_uslot,
                                 const int* rule_ustart, const int* rule_urule, const void* rule_ucoeff,
                                 int n_rule_unique, void* rule_scratch, int slots_disjoint, const void* record,
                                 const void* pull_src,
                                 void* pull_dst, const uint8_t* kind, const void* wall_u, const void* force_field,
                                 const void* body_field, int nx, int ny, int nz, double omega_plus,
                                 double omega_minus, int force_mode, int mask_mode, void* stream);

/* Whether lbm_adj_fused_step takes this problem: float32 D2Q9, bx * by in 64 .. 1024, and an
 * IBM adjoint that stages its tables and transposed rows in LDS (*ok = 1). No counterpart in
 * the reference. Host only. */
LBM_API int lbm_adj_fused_supported(int lattice, int dtype, int n_nodes, int n_markers, int iterations, int n_tslots,
                                    int bx, int by, int* ok);

/* One backward step of the checkpointed adjoint in one launch (no counterpart in the
 * reference): lbm_adj_pull_step (force mode dense, with kind bytes) from lam_next into lam,
 * the IBM adjoint of lbm_ibm_adjoint (same arguments, no rules here; scratch not needed), the
 * transposed link rules of lbm_adj_rules and, when pull_src is not null, an independent
 * forward step from pull_src to pull_dst with the link rules evaluated inside it (as
 * lbm_pull_step_links, force field pull_field) that also writes each rule's value into the
 * slot of pull_src the rule pre-pass would have written. Each lam slot (k, y) of a fluid node
 * has one writer: the push of node y + c_k when that node is fluid, else y's own bounce-back.
 * `region` lists the nodes whose adjoint step the IBM workgroup runs itself (the writers of
 * the transposed moment slots: the IBM stencil nodes). keep_pushes = 1 declares that the
 * region is the IBM node list itself (n_region = n_nodes <= bx * by) and that the transposed
 * slots are exactly the pushes of those nodes, slot (i, x - c_i) listing the moment rows of
 * node x only (density with coefficient 1, then momentum a with c_i,a where nonzero, in row
 * order); the IBM workgroup then keeps the pushes and nu in registers, adds the moment rows
 * itself and neither reads nor stages tslot, tstart, trow and tcoeff. The link rules come as
 * arule_nodes rule nodes (arule_info 16-byte aligned): a rule set in which the target and
 * every term slot of a rule are written by one node, the rule's node, and no term slot is a
 * target; arule_info[4 k] = node (sorted), arule_info[4 k + 1] = the directions i whose
 * upstream node is solid | the directions j whose push writes one of the node's rule targets
 * << 16, the rest zero; arule_coeff[k][i][j] (q x q) = the coefficient of the rule with
 * target push j in the transposed row of the slot push i writes, zero where none, with each
 * transposed row's entries in increasing j. The IBM region and the rule nodes must not meet
 * and their kind bytes carry bit 16 (the other workgroups skip them). next_f, next_record
 * and next_field (may be null) are the state, record and field of the backward step that
 * follows (the preceding time step): one tile workgroup that finishes early reads what that
 * step's IBM workgroup will read first, so it comes from the L2 cache. Same bits as
 * lbm_adj_pull_step_rules followed by lbm_ibm_adjoint_pull. One workgroup shape bx x by for
 * every part. */
LBM_API int lbm_adj_fused_step(
    int lattice, int dtype, const void* f, const int64_t* slot, const void* coeff, const void* constant, int m,
    const int64_t* nodes, int n_nodes, const int* inverse, const void* weights, const void* areas, int n_markers,
    const int* spread_start, const int* spread_entry, int iterations, double fx, double fy, double fz, int64_t n_total,
    void* nu, const double* body_weight, void* lam, const int64_t* tslot, const int* tstart, const int* trow,
    const void* tcoeff, int n_tslots, double* grad_w, const void* record, const void* lam_next, const uint8_t* kind,
    const void* wall_u, const void* force_field, const void* body_field, int nx, int ny, int nz, double omega_plus,
    double omega_minus, double* grad_body, const int32_t* region, int n_region, int keep_pushes,
    const uint32_t* arule_info, const void* arule_coeff, int arule_nodes, int bx, int by, const void* pull_src,
    void* pull_dst, const void* pull_field, const uint32_t* rule_node_info, const int32_t* rule_node_first,
    int rule_nodes, const int64_t* prule_slot, const void* prule_coeff, const void* prule_constant, int64_t n_prules,
    int prule_m, const void* next_f, const void* next_record, const void* next_field, void* stream);

/* Size in bytes of one lattice descriptor of the batched step (no counterpart in the
 * reference, which runs one lattice per solver). Host only. */
LBM_API int lbm_batched_desc_size(int dtype, int64_t* bytes);

/* Fill n_lattices descriptors (host memory, lbm_batched_desc_size bytes each) for lattices
 * of sizes nx/ny/nz[l], rates omega_plus/minus[l] and uniform force force[3 l .. 3 l + 2],
 * laid out back to back in the pools: kind bytes n_l, wall d n_l, and populations q s_l,
 * where population plane i of lattice l starts s_l elements after plane i - 1 and s_l is
 * n_l rounded up to a multiple of 64 (so every plane starts on a cache-line boundary; the
 * padding is never read). No counterpart in the reference (batching is new). */
LBM_API int lbm_batched_describe(int lattice, int dtype, int n_lattices, const int* nx, const int* ny,
                                 const int* nz, const double* omega_plus, const double* omega_minus,
                                 const double* force, void* host_desc);

/* One two-lattice pull step of every lattice of a batch in one launch: `groups` workgroups
 * of bx x by threads walk the n_tiles work items tiles[2 t], tiles[2 t + 1] = (lattice,
 * first node) with a static stride; a work item covers bx * by consecutive node indices
 * of its lattice (rows wrap inside it). desc is the device copy of the descriptors. Same
 * per-node arithmetic as lbm_pull_step (uniform force mode); groups_per_cu caps resident
 * workgroups per compute unit as for the step entry points. No counterpart in the
 * reference. */
LBM_API int lbm_batched_pull_step(int lattice, int dtype, const void* src_pool, void* dst_pool,
                                  const uint8_t* kind_pool, const void* wall_pool, const void* desc,
                                  const int* tiles, int n_tiles, int mask_mode, int bx, int by, int groups,
                                  int groups_per_cu, void* stream);

/* Per-kernel resources of the compiled step kernels for (lattice, dtype):
 * fills up to max_kernels entries of {registers, scratch (private) bytes per lane,
 * static LDS bytes, max threads per block} and writes their names (';'-separated)
 * into names. Needs a device.
 * Replaces CudaSolver.kernel_resources (src/lbm/backends/cuda_solver.py:311-325). */
LBM_API int lbm_kernel_resources(int lattice, int dtype, int* values, int max_kernels,
                                 char* names, int names_len, int* count);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
