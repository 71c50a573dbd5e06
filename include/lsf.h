/*
 * lsf.h -- C ABI of the MI355X (gfx950) FDTD time-stepping library (lightsmith_amd/_lsf.so).
 *
 * The library replaces the per-step arrays operations of lightsmith's FDTD Engine
 * (src/lightsmith/backends/_fdtd/engine.py) and of the per-step source hooks. The reference
 * is pure Python: its "FFI" is the Engine constructor plus the source hook protocol
 * bind/correct_H/advance_H/correct_E/advance_E (engine.py:12-13, 665-678). Each entry point
 * below names the reference interface it stands in for.
 *
 * Conventions
 *   - Every pointer argument named dev_* is device memory owned by the caller (the Python host
 *     allocates it with torch); the library never allocates or frees field, state or table memory.
 *   - Field layout: component-major, then z, y, x with x fastest: element (c, i, j, k) of the
 *     reference's E[3, N0, N1, N2] sits at c*N + (k*N1 + j)*N0 + i, N = N0*N1*N2.
 *     Field type: float (real fields) or complex64 {float re, im} (fields under a Bloch phase).
 *   - Flat indices passed to lsf_add_* are in this native layout (the Python host remaps the
 *     reference's z-fastest indices, engine.py:103-105).
 *   - Every function returns 0 on success, or a nonzero LSF_E* code; lsf_last_error() gives the
 *     message (thread-local). Work is enqueued on the caller's HIP stream; nothing synchronises
 *     the host except lsf_graph_reset.
 *   - Step counter: dev_counter[0] = completed steps (engine.py:271 "n"); dev_counter[1] = the step
 *     in progress, valid for work after the E update (the E kernel copies and increments).
 */
#ifndef LSF_H
#define LSF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSF_OK 0
#define LSF_EINVAL 1   /* bad configuration or argument -> ValueError */
#define LSF_EHIP 2     /* HIP runtime error -> RuntimeError */
#define LSF_ESTATE 3   /* call out of order -> RuntimeError */

typedef struct lsf_engine lsf_engine;

/* Grid description. Replaces Engine.__init__'s scalar arguments (engine.py:231-300). */
typedef struct lsf_grid {
  int32_t shape[3];          /* N0, N1, N2: the reference's axis order (x, y, z) */
  int32_t complex_fields;    /* 1: complex64 fields (any Bloch k != 0, engine.py:257), else float */
  int32_t periodic[3];       /* periodic axes carry ghost planes 0 and n-1 (engine.py:47-66) */
  double spacing[3];         /* cell size per axis (nm) */
  double dt;                 /* time step (nm/c) */
  double bloch_phase[3][2];  /* exp(+i k_a (n_a - 2) d_a) as (cos, sin) per periodic axis (engine.py:595-611) */
  void* dev_E;               /* 3*N field elements */
  void* dev_H;
  int64_t* dev_counter;      /* int64[2], see above */
  int32_t tick;              /* 1: the E kernel advances dev_counter (the top-level engine); 0: a sub-engine */
} lsf_grid;

/* Create / destroy an engine handle (engine.py:230-315; release engine.py:711-719). 3 * cells < 2^31. */
int lsf_create(const lsf_grid* grid, lsf_engine** out);
void lsf_destroy(lsf_engine* eng);

/* Material, dense form: dev_cb[3][N] float = (dt/dx) * inv_eps per E component (engine.py:299), or
 * -(m + 1) for an E node integrated in D-form by pure ADE medium m (engine.py:722-793). */
int lsf_set_material_dense(lsf_engine* eng, const float* dev_cb);

/* Material, indexed form: dev_ids[N] uint32 holding one 8-bit table index per component in bytes 0..2,
 * dev_table[256] float in the dense form's encoding. Same arithmetic as the dense form, 4 B per cell. */
int lsf_set_material_indexed(lsf_engine* eng, const uint32_t* dev_ids, const float* dev_table);

/* CPML on one axis for one field (field 0 = E, 1 = H): replaces Engine._build_pml and the per-slab terms
 * applied in _update_H/_update_E (engine.py:344-418, 636-643, 656-663).
 * dev_profile[shape[axis]] float4 = {b, c, 1/kappa, slab index (or -1)} at each node position along the axis;
 * dev_psi[c] (c != axis) = the psi array of component c, slab extent n_slab along the axis, full extent
 * along the other two (field type). has_kappa = 0 skips the kappa term as the reference does. */
int lsf_set_pml(lsf_engine* eng, int field, int axis, const void* dev_profile, int n_slab, void* dev_psi0,
                void* dev_psi1, void* dev_psi2, int has_kappa);

/* Pure ADE groups in D-form (engine.py:722-742, 759-793). n_nodes D-form nodes in native flat order;
 * dev_chunk_base[(c*N2 + k)*N1 + j][x/32] = index of the 32-cell chunk's first D-form node
 * (ceil(N0/32) chunks per row).
 * 1 <= max_poles <= 4; dev_P[n_nodes][max_poles] pairs {P^n, P^{n-1}} (field type x2, a node's poles adjacent),
 * zeroed by the caller.
 * D is not stored: D^n = eps_inf E^n + sum P^n holds after every step, so the update advances E by
 * (dE - sum (P^{n+1} - P^n)) / eps_inf. dev_media[n_media][2 + 3*max_poles] float =
 * {1/eps_inf, n_poles, (c1, c2, c3) per pole} (pole_coefficients, engine.py:108-111), zeros for the poles past a
 * medium's n_poles (the E kernel advances max_poles poles for every node; zero coefficients keep those poles at
 * zero); n_media * (2 + 3*max_poles) <= 224 (16 media of 4 poles, 44 of 1). */
int lsf_set_ade(lsf_engine* eng, const int32_t* dev_chunk_base, void* dev_P, int64_t n_nodes, int max_poles,
                const float* dev_media, int n_media);

/* Phases of a step at which sparse work runs, in this order (engine.py:665-678). */
#define LSF_PH_CORRECT_H 0  /* after the H update: src.correct_H */
#define LSF_PH_ADVANCE_H 1  /* src.advance_H */
#define LSF_PH_OFFDIAG 2    /* inside _update_E: the off-diagonal gather (engine.py:653-655) */
#define LSF_PH_CORRECT_E 3  /* src.correct_E */
#define LSF_PH_GROUPS 4     /* groups.step for interface (mixed) groups */
#define LSF_PH_ADVANCE_E 5  /* src.advance_E */
#define LSF_PH_BEFORE_H 6   /* before the H update (adjoint engines: the transposed D-form update) */

/* Gather-add rows: F[target[r]] += sum_k w[r, k] * src[src_idx[r, k]] (field 0 = E, 1 = H).
 * Replaces the off-diagonal rows (engine.py:653-655), the 1D-fed TFSF corrections (tfsf.py:195-204)
 * and the Bloch plane corrections (tfsf.py:274-286). dev_scale (or NULL = 1) scales each row's sum: 1/eps_inf
 * for a pure D-form E node, where the reference's correction enters through D before the group update.
 * dev_src NULL reads the engine's own H (its current buffer, see lsf_set_fused). */
int lsf_add_gather(lsf_engine* eng, int phase, int field, int64_t n_rows, int width, const int64_t* dev_target,
                   const void* dev_src, const int64_t* dev_src_idx, const void* dev_w, const float* dev_scale);

/* Modal injection: F[target[r]] += sum_q m[r, q] * table[t * Q + q], t = the step counter, 0 past table_len.
 * Replaces PointDipole.correct_E (emission.py:139-140), ModeSource.correct_H/E (ports.py:368-372) and the
 * Bloch column's current sheet (tfsf.py:291-296). dev_m has the field type; dev_table is float. */
int lsf_add_modal(lsf_engine* eng, int phase, int field, int64_t n_rows, int n_modes, const int64_t* dev_target,
                  const void* dev_m, const float* dev_table, int64_t table_len, const float* dev_scale);

/* Table-fed TFSF corrections (layered.py:169-187, 231-242): for target row r, sum over its points p in
 * [row_ptr[r], row_ptr[r+1]) of coef[p] * quintic(T, t + i0[p], u0[p], base[p]) with tables sampled every
 * m steps, read index clamped to cmax. */
int lsf_add_table(lsf_engine* eng, int phase, int field, int64_t n_rows, const int64_t* dev_target,
                  const int32_t* dev_row_ptr, const float* dev_T, int m, int64_t cmax, const int64_t* dev_base,
                  const int64_t* dev_i0, const float* dev_u0, const float* dev_coef, const float* dev_scale);

/* Incident-wave view of a table source (layered.py:481-482): out[r] = quintic(T, t + i0[r], ...), r < n.
 * Its readers are monitors, so it is evaluated when the engine samples them (lsf_sample, the end of each
 * chunk), with the step counters of the E half of the step just taken; phase must be an E-side phase
 * (LSF_PH_OFFDIAG .. LSF_PH_ADVANCE_E). */
int lsf_add_table_view(lsf_engine* eng, int phase, int n, float* dev_out, const float* dev_T, int m, int64_t cmax,
                       const int64_t* dev_base, const int64_t* dev_i0, const float* dev_u0);

/* Auxiliary 1D Yee grid of a TFSF plane wave (tfsf.py:200-211), float arrays of length M:
 * advance_H: H[m] = da[m] H[m] - db[m] (E[m+1] - E[m]), m < M-1;
 * advance_E: E[m] = ca[m] E[m] - cb[m] (H[m] - H[m-1]), 0 < m < M-1; E[m_src] += wave[t]. */
int lsf_add_line1d(lsf_engine* eng, int M, float* dev_E1, float* dev_H1, const float* dev_ca, const float* dev_cb,
                   const float* dev_da, const float* dev_db, int m_src, const float* dev_wave, int64_t wave_len);

/* A sub-engine (the Bloch plane source's auxiliary column, tfsf.py:256-262, 288-293) whose H and E
 * updates run at LSF_PH_ADVANCE_H and LSF_PH_ADVANCE_E of this engine, on its step counter. */
int lsf_add_subengine(lsf_engine* eng, lsf_engine* sub);

/* Mixed (subpixel two-path) ADE group (engine.py:743-791), at LSF_PH_GROUPS. n nodes, S media slots.
 * dev_idx[n] native flat E indices; dev_f[S][n] fractions; dev_w[n]; dev_h0[n]; dev_inv_a[n];
 * dev_state: field-type state of (3 + S + 4*S*max_poles) * n elements, zeroed by the caller;
 * dev_slot_media[S][2 + 3*max_poles] in lsf_set_ade's media encoding (n_poles = 0: non-dispersive). */
int lsf_add_mixed_group(lsf_engine* eng, int64_t n, int S, int max_poles, const int64_t* dev_idx, const float* dev_f,
                        const float* dev_w, const float* dev_h0, const float* dev_inv_a, const float* dev_slot_media,
                        void* dev_state, int has_h, int has_a);

/* Running DFT monitor (engine.py:556-592): acc_e[f, p] += exp(i w_f t) E_src[e_idx[p]],
 * acc_h[f, p] += exp(i w_f (t - dt/2)) (H_src[h1[p]] + H_src[h2[p]]) / 2, t = dev_counter[0] * dt, phases in
 * double and cast to complex64. energy_group 1 or 2 adds sum |E|^2 + |H|^2 into that energy slot
 * (engine.py:580-589); n_freq <= 2048. src pointers default to this engine's fields when NULL (a source's own arrays
 * otherwise: src_complex gives their type). */
int lsf_add_dft_monitor(lsf_engine* eng, int n_freq, int64_t n_pts, const double* dev_omega,
                        const int64_t* dev_e_idx, const int64_t* dev_h1, const int64_t* dev_h2, void* dev_acc_e,
                        void* dev_acc_h, const void* dev_e_src, const void* dev_h_src, int src_complex,
                        int energy_group);

/* Time monitor (engine.py:213-224, 590-592): buf[pos, p] = E[e_idx[p]], pos += 1 per sample, pos < n_max. */
int lsf_add_time_monitor(lsf_engine* eng, int64_t n_pts, const int64_t* dev_e_idx, void* dev_buf, int64_t n_max,
                         int64_t* dev_pos);

/* Conformal vertex tensor (TensorMesh, engine.py:140-181, 796-923): the D-form update of n_nodes interface E
 * nodes (native flat indices dev_nodes, components dev_comp) from n_vert tensor vertices. dev_vert_nodes[v][d][2]:
 * node numbers of vertex v's minus/plus d-nodes; dev_vert_phase (complex64 [v][d][2], complex fields only):
 * exp(i k . shift L) of those node positions; dev_normal[v][3]. Slots q = k * n_nodes + node (k = 0: the vertex at
 * the node's index, 1: at index + e_c): dev_slot_vert[q] = tensor vertex or -1 (plain slot). Per slot the filter
 * inputs dev_frac[S][2 n_nodes] (medium fractions; one-hot for plain slots), dev_h0 (sum of f_s / eps_inf_s over
 * non-dispersive media), dev_inv_a (1 / sum f_s eps_inf_s); dev_slot_media[S][2 + 3 max_poles] = {1/eps_inf,
 * n_poles, (c1, c2, c3) per pole}. dev_state (field type, zeroed by the caller):
 * D[n] | E_prev[n] | out[2n] | Ea[2n] | Eh[S][2n] | Ph[S][max_poles][2][2n] | Pa[S][max_poles][2][2n], n = n_nodes.
 * Runs after the E update and the E corrections, with the other groups (engine.py:672-676). */
int lsf_add_tensor_group(lsf_engine* eng, int64_t n_nodes, const int64_t* dev_nodes, const int32_t* dev_comp,
                         int64_t n_vert, const int32_t* dev_vert_nodes, const void* dev_vert_phase,
                         const float* dev_normal, const int32_t* dev_slot_vert, int S, int max_poles,
                         const float* dev_frac, const float* dev_h0, const float* dev_inv_a,
                         const float* dev_slot_media, void* dev_state);

/* Launch the tensor group's D update (D += E - E_prev per node) with the E-side table correction, as one kernel:
 * the last op of the E correction phase must be a table correction on E and the first op of the group phase a
 * tensor group. The caller guarantees that no table target is a tensor node, so the D update reads the same E as
 * after a separate correction launch. No counterpart in the reference (it fuses nothing); adding an E correction
 * or group op afterwards undoes it. */
int lsf_fuse_table_tensor(lsf_engine* eng);

/* Energy proxy slots (engine.py:315 "energy"), float[2] (or a field-precision pair). */
int lsf_set_energy(lsf_engine* eng, float* dev_energy);

/* Launch tuning: rows per lane (1 or 2), wavefronts per block (4 or 8), x cells per wavefront row (64: one
 * row per wavefront; 32: two rows), z planes per block (>= 1; the E kernel of an engine with pole state marches
 * twice as many). No reference counterpart (a launch choice). */
int lsf_set_tiling(lsf_engine* eng, int rows_per_lane, int waves, int x_cells, int z_chunk);

/* Order of the E launch's tiles: reverse_e = 1 (the default) walks each XCD's contiguous tile range from its end,
 * so the E kernel starts on the tiles whose data the H kernel touched last and the next H kernel on those the E
 * kernel touched last (the 256 MB memory-side cache still holds part of them); 0 walks forward like the H launch.
 * Results do not depend on it (tiles are independent within a launch). No counterpart in the reference. */
int lsf_set_tile_order(lsf_engine* eng, int reverse_e);

/* Run n_steps steps (Engine.step, engine.py:665-678) on `stream` (a hipStream_t; NULL = default). */
int lsf_step(lsf_engine* eng, int n_steps, void* stream);
/* Run n_steps steps as lsf_step does, allocating nothing. No counterpart in the reference: for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc). The two kernel-argument blocks
 * that lsf_step keeps in a library allocation are staged instead, on `stream`, in one file-scope device array of
 * 8 KiB per GPU, shared by every call on that GPU: calls that overlap on one GPU (other streams or host threads)
 * race. Refused with an error: engines with sub-engines (their arguments exceed the array) and single-pass
 * engines. */
int lsf_step_noalloc(lsf_engine* eng, int n_steps, void* stream);
/* As lsf_step, bracketing each H and E kernel with HIP events on `stream`; adds the summed kernel times (ms)
 * to ms[0] (H kernels) and ms[1] (E kernels; the single-pass kernel when the engine steps with it, ms[0] then
 * gaining 0). Synchronises the stream at the end. For measurement. */
int lsf_step_timed(lsf_engine* eng, int n_steps, void* stream, double* ms);
/* One DFT/time-monitor sample (Engine._sample, engine.py:568-592). */
int lsf_sample(lsf_engine* eng, void* stream);
/* n_chunks x (dft_stride steps + one sample) (Engine.chunk, engine.py:680-683). */
int lsf_chunk(lsf_engine* eng, int n_chunks, int stride, void* stream);

/* HIP graph of one chunk (make_runner, engine.py:685-709): capture on `stream` (not the legacy default
 * stream), replay n times, drop. */
int lsf_graph_capture(lsf_engine* eng, int stride, void* stream);
int lsf_graph_replay(lsf_engine* eng, int n, void* stream);
int lsf_graph_reset(lsf_engine* eng);

/* ---- Adjoint (gradient) support -------------------------------------------------------------------
 * No reference counterpart: the reference FDTD has no adjoint (its inverse design runs on FDFD). An adjoint
 * engine is an engine created like its forward twin (same grid, material and CPML profiles; its own field,
 * counter and CPML state arrays, which hold the adjoint variables e~ = cb dL/dE, h~ = -cH dL/dH and the
 * transposed psi terms) and switched to the transposed, time-reversed update. Each lsf_step of it reverses one
 * forward step, last step first. Real fields, no periodic axes, no D-form nodes. */
int lsf_set_adjoint(lsf_engine* eng, int on);
/* Forward tape: during the H phase of each forward step n, record X = curl- H^{n+1/2} (the term cb multiplies
 * in the E update) at n_nodes design E nodes (native flat indices, interior) into dev_tape[n - base][node]
 * (float), for 0 <= n - dev_base[0] < n_slots. */
int lsf_add_tape(lsf_engine* eng, int64_t n_nodes, const int64_t* dev_nodes, float* dev_tape, int64_t n_slots,
                 const int64_t* dev_base);
/* Adjoint engine: after each adjoint H half-step reversing forward step n = n_steps - 1 - steps done,
 * dev_g[q] += (e~/cb)[node q] * tape[n - base][q] (double), i.e. dL/dcb at each design node. */
int lsf_add_adjoint_gradient(lsf_engine* eng, int64_t n_nodes, const int64_t* dev_nodes, const float* dev_cb,
                             const float* dev_tape, int64_t n_slots, const int64_t* dev_base, int64_t n_steps,
                             double* dev_g);
/* Adjoint engine: the transposed D-form (ADE) update of a forward engine's pure-group nodes (engine.py:759-793
 * run backwards): n nodes (native flat E indices, sorted), their medium, and the media table of the forward engine
 * (n_media x {1/eps_inf, poles, (c1, c2, c3) x max_poles}); state holds the adjoint pole pairs (n x max_poles x 2
 * floats, zeroed by the caller), save n floats of working space. Before each adjoint H kernel the nodes' adjoint
 * fields become e~ / eps_inf (what the transposed curl reads) and the pole recursion advances; after it the new
 * adjoint fields are put back. No reference counterpart (the reference has no adjoint). */
int lsf_add_adjoint_ade(lsf_engine* eng, int64_t n, const int64_t* dev_nodes, const int32_t* dev_medium, int n_media,
                        int max_poles, const float* dev_media, float* dev_state, float* dev_save);
/* Adjoint engine: the source of a DFT monitor objective. dev_W[nf][n_pts] complex64 = conj(dL/dacc) (torch's
 * gradient of the monitor's E or H accumulator). The targets are given as CSR rows, one per distinct node:
 * row r adds to node dev_target[r] the sum over its entries q (dev_row_ptr[r] .. dev_row_ptr[r+1]) of
 * dev_scale[q] * sum_f Re(W[f, dev_point[q]] exp(i omega_f t)), summed in that fixed order (deterministic where
 * monitor points share a node). E side: after the adjoint E half-step reaching E^n, for n > 0, n % stride == 0,
 * t = n dt (scale = cb at the node). H side: after the adjoint H half-step reaching H^{n-1/2}, t = n dt - dt/2;
 * each H point enters the rows of both its averaged nodes (scale = -cH / 2). The source of E^{n_steps} itself is
 * the caller's initial condition. */
int lsf_add_adjoint_source(lsf_engine* eng, int h_side, int nf, int64_t n_pts, const double* dev_omega, double dt,
                           const void* dev_W, int64_t n_rows, const int32_t* dev_row_ptr, const int64_t* dev_target,
                           const int64_t* dev_point, const float* dev_scale, int64_t n_steps, int stride);

/* ---- Batches: many independent engines stepped together -----------------------------------------
 * No reference counterpart: the reference runs the angles/wavelengths/polarisations of an oblique sweep
 * one Engine after another (periodic.py:503-519). A batch steps its members with shared launches: per
 * half-step one bulk-kernel launch per kernel configuration over all members' tiles (one block per tile,
 * or a persistent grid that strides over them), each phase's per-step operations fused by position over
 * the members (gathers and modal sources one launch each, sub-engines as a nested batch), and one DFT
 * launch per source type for all monitors. Members are fully configured top-level engines with their own
 * step counters; while the batch exists they refuse configuration calls. Destroy the batch before its
 * members. Each member's state evolves exactly as under lsf_step / lsf_chunk of that member alone. */
typedef struct lsf_batch lsf_batch;
int lsf_batch_create(lsf_engine* const* members, int n_members, lsf_batch** out);
void lsf_batch_destroy(lsf_batch* batch);
/* Tile shape as lsf_set_tiling, for all members' bulk kernels; grid_blocks = 0 launches one block per tile,
 * grid_blocks > 0 a persistent grid of that many blocks (rounded up to a multiple of 8). */
int lsf_batch_set_schedule(lsf_batch* batch, int rows_per_lane, int waves, int x_cells, int z_chunk,
                           int grid_blocks);
/* Tiles of one E half-step over all members under the current schedule. */
int lsf_batch_tiles(lsf_batch* batch, int64_t* n_tiles);
int lsf_batch_step(lsf_batch* batch, int n_steps, void* stream);
/* As lsf_batch_step, adding the summed bulk-kernel times (ms) to ms[0] (H) and ms[1] (E). Synchronises. */
int lsf_batch_step_timed(lsf_batch* batch, int n_steps, void* stream, double* ms);
int lsf_batch_sample(lsf_batch* batch, void* stream);
int lsf_batch_chunk(lsf_batch* batch, int n_chunks, int stride, void* stream);
int lsf_batch_graph_capture(lsf_batch* batch, int stride, void* stream);
int lsf_batch_graph_replay(lsf_batch* batch, int n, void* stream);
int lsf_batch_graph_reset(lsf_batch* batch);

/* Single-pass stepping (no reference counterpart: a launch choice). dev_E2, dev_H2: second field buffers (zeroed,
 * like dev_E/dev_H); dev_psiH2[axis]: a second set of the H-side psi arrays of each axis with CPML (null where
 * none), laid out like the first, whose two component arrays must be contiguous (the second component's array
 * right after the first's). Where the engine allows it -- real fields, no periodic axes, one row per lane and 64 x
 * cells per row, no operations between the H and the E update -- each step is then one kernel that reads one buffer
 * set and writes the other; the E-phase operations, monitors and graphs follow the current set. lsf_parity gives
 * the current set (0: dev_E/dev_H, 1: dev_E2/dev_H2). Results are bit-identical to the two-kernel step. */
int lsf_set_fused(lsf_engine* eng, void* dev_E2, void* dev_H2, void* const* dev_psiH2);
int lsf_parity(lsf_engine* eng);

/* Message of the last failing call on this thread. */
const char* lsf_last_error(void);
/* Library version string. */
const char* lsf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LSF_H */
