/* C ABI of libpfc_hip.so: the MI355X (gfx950) pseudo-spectral ETD step of the
 * standard and log phase-field-crystal models.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative PFC_E* code; the
 *     message of the last failure on the calling thread is pfc_last_error().
 *     No C++ exception crosses this boundary.
 *   - Fields are real, C-order (Ny, Nx), x fastest, in DEVICE memory owned by
 *     the caller (e.g. a torch tensor's data_ptr()).  `dtype` is 64 (double)
 *     or 32 (float).  Spectra are (Ny, Nx/2+1) interleaved complex, ky in
 *     numpy.fft.fftfreq order, forward transforms unnormalised, inverses 1/N.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Stepping is
 *     asynchronous on it; pfc_reduce synchronises it.
 *   - A plan owns its device state, tables and scratch; it never frees caller
 *     memory.  A plan is not thread safe: one host thread per plan at a time.
 *
 * Reference interfaces replaced (paths under EricStimpsonWSU/PFC-models):
 *   pfc_create      Build*Simulation + _SimulationState tables
 *                   (src/pfc_models/models/stdpfc/simulation.py:99-164, :270-286;
 *                    src/pfc_models/discretization/spectral/kernels.py:38-71)
 *   pfc_set_field   `sim.State.n[...] = x` (models/stdpfc/simulation.py:148-153)
 *   pfc_get_field   reading `sim.State.n`
 *   pfc_step        Simulation.TimeStep (models/stdpfc/simulation.py:225-235;
 *                   models/logpfc/simulation.py:220-230), nsteps in a row
 *   pfc_reduce      Simulation.CalculateFreeEnergy (models/stdpfc/simulation.py:249-268;
 *                   models/logpfc/simulation.py:244-263) and `State.n.mean()`
 *                   (tests/test_smoke_models.py:245-249)
 *   pfc_set_shift   logpfc_extv OptimizeM table rebuild
 *                   (models/logpfc_extv/simulation.py:332-373)
 *   pfc_rfft2 / pfc_irfft2
 *                   the backend FFT contract `rfftn`/`irfftn`
 *                   (src/pfc_models/backends/torch_mps.py:3-14) as used by
 *                   cp.fft.rfft2 / irfft2 (models/stdpfc/simulation.py:215-223)
 */
#ifndef PFC_HIP_H
#define PFC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PFC_OK 0
#define PFC_EINVAL (-1)       /* bad argument */
#define PFC_EUNSUPPORTED (-2) /* grid size not factorable / too large */
#define PFC_EHIP (-3)         /* HIP runtime error */
#define PFC_ENOMEM (-4)       /* device allocation failed */

#define PFC_MODEL_STD 0 /* field n:   N = g n^2 + n^3 */
#define PFC_MODEL_LOG 1 /* field phi: N = ln phi + g phi^2 + phi^3 - M phi */

#define PFC_SCHEME_ETD1 1
#define PFC_SCHEME_ETDRK2 2

#define PFC_STATUS_NONFINITE 1 /* a non-finite field value was seen by a step */
#define PFC_STATUS_ABORTED 2   /* a persistent group step stopped early (pfc_group_set_flow): the field is undefined */

typedef struct pfc_plan pfc_plan;

/* Model, grid and numerics of one instance.  Field names follow
 * Build2DGeometry / Build{Std,Log}PFCModelParameters. */
typedef struct pfc_params {
  int model;     /* PFC_MODEL_* */
  int Nx, Ny;    /* Nx even; both factorable into 2,3,5,7,11,13 */
  double Lx, Ly; /* box lengths: dx = Lx/Nx, dy = Ly/Ny */
  double Gamma, beta, q0, epsilon, g, dt;
  double minPhi;    /* logPFC clamp floor (used when clamp_log) */
  double smallZTol; /* ETD1 Taylor switch |L dt| <= smallZTol (kernels.py:55-71) */
  double M;         /* logPFC stabilising shift (0 = none) */
  int precision;    /* 64: FP64 bulk; 32: FP32 bulk (transforms, state) with FP64 ETD tables
                       (the fast column kernel reads their FP32 copies, Ny E1 - 1 and Q1,
                       widened to FP64), FP64 combine, FP64 reductions */
  int scheme;       /* PFC_SCHEME_* */
  int clamp_log;    /* logPFC: ln(phi) -> ln(minPhi) where phi <= minPhi */
  int pin_mean;     /* pin the k=0 mode of every update to the mean captured by pfc_set_field */
} pfc_params;

const char* pfc_last_error(void);
int pfc_version(void);

int pfc_create(const pfc_params* params, pfc_plan** out);
int pfc_destroy(pfc_plan* plan);

/* Load a real (Ny, Nx) device field; captures its FP64 mean as the k=0 pin. */
int pfc_set_field(pfc_plan* plan, const void* field, int dtype, void* stream);
/* Write the current real (Ny, Nx) field to device memory. */
int pfc_get_field(pfc_plan* plan, void* field, int dtype, void* stream);
/* Advance nsteps time steps (asynchronous). */
int pfc_step(pfc_plan* plan, long long nsteps, void* stream);
/* FP64 compensated reductions of the current field; synchronises `stream`.
 * out[0] = free energy per site (the box mean of f), out[1] = mean,
 * out[2] = min, out[3] = max. */
int pfc_reduce(pfc_plan* plan, double out[4], void* stream);
/* Tracker: advance `nsteps` steps and record pfc_reduce's four values after
 * every `every` steps into host `out` ((nsteps / every) x 4 doubles), the
 * records written on the device and copied back with one synchronisation at
 * the end (steps past the last full block run unrecorded).  Replaces the
 * TimeStep / CalculateFreeEnergy / State.n.mean() loop of the reference's
 * callers (tests/test_smoke_models.py:104-113; the driver's per-block free
 * energy, phase_diagram/q0star.py:261-307), which synchronises per record. */
int pfc_run_tracked(pfc_plan* plan, long long nsteps, long long every, double* out, void* stream);
/* Rebuild the ETD tables with the stabilising shift M (logPFC). */
int pfc_set_shift(pfc_plan* plan, double M);
/* Stabilising shift managed on the device (logPFC; on = 1): every step, before
 * the shift acts, the library takes the field's min and max (the values the
 * step's inverse row transforms give; pfc_reduce's where no fast row shape
 * exists), and when the slot has no fit yet, or the min fell below the fitted
 * floor lo by more than
 * floor_tol (relative) or the max rose above the fitted ceiling hi by more than
 * ceil_tol, refits lo = (min > 0 ? min : minPhi), hi = max,
 * M = (N(hi) - N(lo)) / (hi - lo), N(phi) = g phi^2 + phi^3 + ln phi, and
 * rebuilds that slot's ETD tables, all on the stream without a host
 * synchronisation.  Replaces logpfc_extv's per-step ValidatePhiRange /
 * OptimizeM (models/logpfc_extv/simulation.py:332-405: two device-to-host
 * copies and a host table rebuild per step; tolerances 0.1 there).  A new
 * field or slot parameters make the next step refit.  Under ETD1 the shift
 * is folded into the ETD table, (E - M P Gamma lap) n_hat + P Gamma lap N_hat
 * with N computed without -M phi, so the check sits between the step's row and
 * column passes (equal to the unfolded update up to rounding).  on = 0 hands M
 * back to pfc_set_shift; on = 2 (the shift already on) replaces the two
 * tolerances and keeps every slot's fit, as changing phi_floor_tol /
 * phi_ceiling_tol between steps does in the reference, which reads them at
 * every check.  Not with groups or per-slot step counts. */
int pfc_set_auto_shift(pfc_plan* plan, int on, double floor_tol, double ceil_tol, void* stream);
/* out = {M, lo, hi} (lo = hi = NaN before the first fit); synchronises. */
int pfc_get_shift(pfc_plan* plan, double out[3], void* stream);
/* Read and clear the PFC_STATUS_* flags (synchronises `stream`). */
int pfc_get_status(pfc_plan* plan, int* flags, void* stream);
/* Chemical potential of the current field, real (Ny, Nx) device outputs:
 * mu = irfft2(L_mu n_hat + N_hat) and local_mu = irfft2(G (L_mu n_hat + N_hat)),
 * G = exp(-k^2 w^2/2), w = 2 pi.  Replaces Batch2 (_Batch2_CalcMu_Terms ->
 * State.mu / State.local_mu, models/stdpfc/simulation.py:217-219, :299-313). */
int pfc_chemical_potential(pfc_plan* plan, void* mu, void* local_mu, int dtype, void* stream);
/* Gaussian-smoothed field irfft2(G n_hat) into a real (Ny, Nx) device output.
 * Replaces Get_local_n / Get_local_phi (models/stdpfc/simulation.py:237-247). */
int pfc_local_field(pfc_plan* plan, void* out, int dtype, void* stream);
/* External potential V, a real (Ny, Nx) device array of `dtype` (NULL removes
 * it): the update becomes n_hat' = E n_hat + P (Gamma lap N_hat + lap V_hat),
 * the forcing of models/logpfc_extv/simulation.py:557-577 (no Gamma on the V
 * term; with ETDRK2 the constant term enters the first stage only).
 * V_hat = rfft2(V) is kept in FP64 and the source P lap V_hat is rebuilt with
 * the ETD tables when dt, epsilon, M or the box change.  The free energy does
 * not include the potential, as the reference's does not
 * (models/logpfc_extv/model.md, "The reported free energy excludes the
 * external potential").  A plan with a potential steps with the generic
 * column kernel, or the resident stepper where that applies.  Synchronises
 * `stream`. */
int pfc_set_external_potential(pfc_plan* plan, const void* V, int dtype, void* stream);
/* Spectral diagnostics of the current field, from the row spectrum on the
 * device (n_hat = rfft2(n), N = Nx Ny, sums over the full spectrum):
 * out[0] = mean = n_hat(0,0)/N, out[1] = sum |n_hat/N|^2 (= mean of n^2),
 * out[2] = sum |n_hat/N|^2 |k|^2, out[3] = sum |n_hat/N|^2 |k|^4, then
 * out[4 + 2e], out[5 + 2e] = n_hat at (ky[e], kx[e]) for e < nent (<= 256;
 * 0 <= ky < Ny, 0 <= kx <= Nx/2).  The phase-diagram adapters' 3-mode
 * fitness and amplitudes (models/stdpfc_cpu/phase_diagram.py:193-205) and the
 * frozen-field moments (models/logpfc_cpu/phase_diagram.py:475-500) follow
 * from these without the field leaving the device.  Synchronises `stream`.
 * PFC_EUNSUPPORTED if two FP64 copies of the half spectrum do not fit in LDS. */
int pfc_diagnose(pfc_plan* plan, const int* ky, const int* kx, int nent, double* out, void* stream);
/* Kernel timing for roofline figures: each kernel kind of the plan's current
 * schedule runs `nsteps` times back to back on scratch copies of the live
 * state (the plan's state is not advanced), bracketed by one HIP event pair on
 * `stream`; ms[k] is the total and calls[k] the launch count of kind k (0 row
 * pass, 1 first column pass, 2 ETDRK2 second column pass; 0 calls when the
 * scheme has no such pass; on the resident stepper kind 0 is one whole step
 * per launch and kinds 1, 2 are empty).  Runs the autotuner first if no step has.
 * Synchronises `stream`.  (No reference counterpart: profiling
 * instrumentation lives outside the model code, README.md:59-60.) */
int pfc_step_profile(pfc_plan* plan, long long nsteps, void* stream, double ms[3], long long calls[3]);
/* Kernel schedule in use (after the first step has run the autotuner):
 * info[0..3] = row kernel  {compile-time schedule in use, variant (+16: the FP32 form held to three
 *               waves per SIMD), points per lane, rows per block},
 * info[4..7] = column kernel {same, columns per block}.  A zero in info[0] or
 * info[4] means the generic mixed-radix kernel runs; -1 in both means the
 * resident stepper runs (small ETD1 instances: one workgroup per instance
 * holds the whole row spectrum in LDS for all the steps of a call), -2 its
 * dense form (cells of at most (Nx/2 + 1) Ny = 512 points, Nx, Ny <= 32;
 * PFC_HIP_RESIDENT_DENSE=0 keeps the FFT form).
 * (No reference counterpart.) */
int pfc_plan_schedule(const pfc_plan* plan, int info[8]);
/* Store kind of the fast row and column kernels' outputs (after the first
 * step): 1 non-temporal, 0 ordinary, -1 when the two fast kernels do not both
 * run.  The autotuner times a step loop with each and keeps the faster;
 * PFC_HIP_STORES=nt|ordinary pins it.  The kind changes no result.
 * (No reference counterpart.) */
int pfc_plan_stores(const pfc_plan* plan, int* nt);
/* Step-loop graph counters: out[0] = graphs captured, out[1] = graph replays.
 * A pfc_step / pfc_batch_step call of n >= 32 steps on the two-kernel path
 * replays the HIP graph of 32 steps of its launch configuration n / 32 times
 * and launches the remaining steps directly; a configuration without a graph
 * gets one captured by the first call of n >= 256 steps (shorter calls launch
 * directly).  Results are bit-identical to direct launches.  PFC_HIP_GRAPH=0
 * disables the graphs.
 * (No reference counterpart: the reference's callers loop TimeStep(),
 * models/stdpfc/simulation.py:213-235, one kernel sequence per step.) */
int pfc_plan_graph_stats(const pfc_plan* plan, long long out[2]);
int pfc_batch_graph_stats(const struct pfc_batch* batch, long long out[2]);
/* Device bytes held by a plan (state, tables, scratch). */
int pfc_plan_bytes(const pfc_plan* plan, long long* bytes);

/* ---- batches: many same-shape instances stepped by the same launches ----
 * A batch holds `capacity` slots sharing the grid, model, scheme, precision
 * and every parameter of pfc_params except epsilon and M (pfc_batch_set_slot)
 * and the box Lx, Ly and time step dt (pfc_batch_set_slot_box), which are per
 * slot, initially the shared values.  Each slot has its own field, k-grids,
 * ETD tables, k = 0 pin and status word.  pfc_batch_step advances the
 * active slots (initially all) together: one row launch and one column launch
 * per nonlinear evaluation, whatever the number of slots.  A slot's results do
 * not depend on which other slots are active.  (No reference counterpart: the
 * reference relaxes one Simulation at a time, phase_diagram/q0star.py:409-477
 * and :566-635; this is the batched submission the phase-diagram driver and a
 * serving loop use.) */
typedef struct pfc_batch pfc_batch;
int pfc_batch_create(const pfc_params* shared, int capacity, pfc_batch** out);
int pfc_batch_destroy(pfc_batch* batch);
/* Per-slot epsilon and stabilising shift M; rebuilds that slot's tables (synchronises `stream`). */
int pfc_batch_set_slot(pfc_batch* batch, int slot, double epsilon, double M, void* stream);
/* Per-slot box Lx x Ly and time step dt: the slot's k-grids (the k-grids of
 * Build2DGeometry, discretization/spectral/geometry_2d.py:70-88) and ETD
 * tables are rebuilt (synchronises `stream`); the field is kept as stored, so
 * set it again when the box changes meaning.  The phase-diagram driver's
 * evaluations at different lattice scales q00 (phase_diagram/q0star.py:409-477)
 * each need their own box, and its rollback refinement its own dt (:261-307). */
int pfc_batch_set_slot_box(pfc_batch* batch, int slot, double Lx, double Ly, double dt, void* stream);
int pfc_batch_set_field(pfc_batch* batch, int slot, const void* field, int dtype, void* stream);
int pfc_batch_get_field(pfc_batch* batch, int slot, void* field, int dtype, void* stream);
/* Fields of the active slots, in active order, into a device array
 * (n_active, Ny, Nx) of `dtype` (one launch for all of them).  The batched
 * phase-diagram driver reads every running evaluation's field per block
 * (phase_diagram/q0star.py:261-307 reads one with get_state). */
int pfc_batch_get_fields(pfc_batch* batch, void* fields, int dtype, void* stream);
/* pfc_set_external_potential for one slot of a batch. */
int pfc_batch_set_external_potential(pfc_batch* batch, int slot, const void* V, int dtype, void* stream);
/* pfc_diagnose for every active slot: out[n_active][4 + 2 nent], active order. */
int pfc_batch_diagnose(pfc_batch* batch, const int* ky, const int* kx, int nent, double* out, void* stream);
/* Slots stepped and reduced by the next calls, in this order (synchronises `stream`). */
int pfc_batch_set_active(pfc_batch* batch, const int* slots, int n, void* stream);
int pfc_batch_step(pfc_batch* batch, long long nsteps, void* stream);
/* counts[i] ETD1 steps for the i-th active slot (active order), in one
 * launch per 4096 steps of the largest count: the slots of a batch advance by
 * different numbers of steps, as the phase-diagram driver's adaptive time
 * step (q0star.py:309-326) gives each evaluation its own dt per sub-block.
 * Resident-stepper engines only (small ETD1 grids; PFC_EUNSUPPORTED
 * otherwise).  Synchronises `stream`. */
int pfc_batch_step_counts(pfc_batch* batch, const long long* counts, void* stream);
/* FP64 reductions of the active slots: out[4 i + {0,1,2,3}] = free energy per
 * site, mean, min, max of the i-th active slot.  Synchronises `stream`. */
int pfc_batch_reduce(pfc_batch* batch, double* out, void* stream);
/* pfc_set_auto_shift / pfc_get_shift for every slot of a batch / one slot. */
int pfc_batch_set_auto_shift(pfc_batch* batch, int on, double floor_tol, double ceil_tol, void* stream);
int pfc_batch_get_shift(pfc_batch* batch, int slot, double out[3], void* stream);
/* pfc_run_tracked for the active slots: out[rec][i][4] for the i-th active slot. */
int pfc_batch_run_tracked(pfc_batch* batch, long long nsteps, long long every, double* out, void* stream);
/* Read and clear every slot's PFC_STATUS_* flags: flags[capacity]. */
int pfc_batch_get_status(pfc_batch* batch, int* flags, void* stream);
int pfc_batch_bytes(const pfc_batch* batch, long long* bytes);
/* Device bytes one slot of these parameters needs (state + tables, FP32
 * engines' FP32 table copies included); for admission control. */
int pfc_slot_bytes(const pfc_params* params, long long* bytes);

/* ---- groups: batches of different grid sizes advanced by shared launches ----
 * A group refers to (does not own) `n` batches that share precision and
 * scheme; each has Nx/2 and Ny in {32, 48, 64, 96, 128, 192, 256, 384, 512},
 * is not small enough for the resident stepper and has no external potential.
 * pfc_group_step advances the active slots of every member by `nsteps`
 * steps with, per nonlinear evaluation, one row launch and one column launch
 * for each register class that has members (class 0: lengths 32, 64, 128,
 * 256, 512; class 1: 48, 96, 192, 384), whatever the number of sizes: the
 * launches' blocks are the members' row (column) tiles, largest instances
 * first, each block finding its member, instance and tile from the members'
 * block offsets.  The transforms are the variant-0
 * schedules, so a slot's result equals its batch stepped alone with that
 * schedule (PFC_HIP_TUNE=0), bit for bit.  Calls of >= 32 steps replay
 * step-loop graphs as pfc_batch_step does.  The batches stay usable on their
 * own; a member must outlive the group, and members' active lists may change
 * between calls (the next call rebuilds the launch layout after draining
 * `stream`).  With PFC_HIP_GROUP_CACHE_MB set to a budget in megabytes
 * (default 0: one chunk), members are stepped in chunks, largest first,
 * whose state arrays and tables fit it: a step graph runs 32 steps of one
 * chunk, then of the next, so a chunk's working set can stay in the 256 MB
 * Infinity Cache across its steps; chunks are independent, so the chunking
 * changes no result.  pfc_group_chunks gives the number of chunks of the
 * last call.  The launches' output stores are ordinary ones
 * (PFC_HIP_GROUP_STORES=nt: non-temporal).
 * (No reference counterpart: the reference relaxes one Simulation
 * at a time, phase_diagram/q0star.py:409-477; this is the uneven-size sweep
 * form, one launch per kernel kind per step over a work list of FFT tiles.) */
typedef struct pfc_group pfc_group;
int pfc_group_create(pfc_batch* const* batches, int n, pfc_group** out);
int pfc_group_destroy(pfc_group* group);
int pfc_group_step(pfc_group* group, long long nsteps, void* stream);
int pfc_group_graph_stats(const pfc_group* group, long long out[2]);
int pfc_group_chunks(const pfc_group* group, int* nchunks);
/* Persistent form of pfc_group_step.  steps_per_launch > 0: each register
 * class's members advance in one launch per `steps_per_launch` steps
 * (k_flow), whose blocks take the tiles of a work list in order -- for each
 * step, each phase (rows, columns; ETDRK2: rows, stage 1, rows, stage 2),
 * each member (largest first), each active instance, each tile -- and start
 * a tile once its instance has finished every tile of the phases before it
 * (per-instance counters), so instances' passes overlap with no launch
 * boundary or device-wide drain between passes or steps; the two classes'
 * launches run side by side.  Tiles run the same transforms as the
 * per-pass launches, so results are bit-identical to them.  Progress does
 * not depend on how many blocks are resident.  A dependency wait that
 * exceeds its bound (~1 s; not expected; PFC_HIP_FLOW_SPIN sets the number
 * of polls) stops the launch and flags every member slot PFC_STATUS_ABORTED.  Needs each member's Nx/2 and Ny in one
 * register class (PFC_EUNSUPPORTED otherwise).  0: the per-pass launches.
 * pfc_group_flow_stats: out[0] launches so far, out[1..2] blocks per launch
 * of classes 0 and 1 (resident capacity), out[3] steps per launch.
 * (No reference counterpart: the reference relaxes one Simulation at a time,
 * phase_diagram/q0star.py:409-477; this is the persistent uneven-size form of
 * the sweep.) */
int pfc_group_set_flow(pfc_group* group, int steps_per_launch);
int pfc_group_flow_stats(const pfc_group* group, long long out[4]);

/* Batched 2D real transforms over (batch, Ny, Nx) device arrays.
 * pfc_rfft2:  out[b] = rfft2(in[b])             (unnormalised)
 * pfc_irfft2: out[b] = irfft2(in[b], s=(Ny,Nx)) (scaled 1/(Nx Ny); `in` is not modified).
 * pfc_irfft2 transforms a working copy of `in` held per (device, stream), so
 * calls on different streams may overlap; it grows (hipMalloc) when a call
 * needs more than the stream's copy holds. */
int pfc_rfft2(const void* in, void* out, int batch, int Ny, int Nx, int dtype, void* stream);
int pfc_irfft2(const void* in, void* out, int batch, int Ny, int Nx, int dtype, void* stream);
/* pfc_irfft2 without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as pfc_irfft2 (its roots are computed on the device,
 * so results may differ from pfc_irfft2's in the last bit of a root).  The
 * roots and the working copy of the spectrum live in one static device array
 * of 64 MiB per device, held until the process ends:
 * batch*Ny*(Nx/2+1)*(16 or 8) bytes plus the roots must fit in it, else
 * PFC_EINVAL.  Every call on a device shares that array, so calls that
 * overlap on one device (different streams, or threads) race; calls on one
 * stream are ordered. */
int pfc_irfft2_noalloc(const void* in, void* out, int batch, int Ny, int Nx, int dtype, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PFC_HIP_H */
