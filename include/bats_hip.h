/*
 * bats_hip.h -- C ABI of the MI355X (gfx950) implementation of the BATS-R-US
 * explicit block-adaptive ideal-MHD stage update.
 *
 * The reference's hot path is a set of Fortran module procedures working on
 * module-global block arrays (State_VGB, LeftState_V*, Flux_V*I, DtMax_CB);
 * it has no C or FFI layer of its own.  Each entry point below replaces one of
 * those procedures and keeps its name, argument meaning and order, so that a
 * Fortran host can bind them with ISO_C_BINDING (see the bind(C) module
 * batsrus_amd/fortran/ModHipBridge.f90).  Differences from the reference:
 *   - the arrays live on the GPU inside an opaque handle; the host exchanges
 *     the reference's State_VGB layout only at sync points (upload/download),
 *     mirroring sync_cpu_gpu (src/ModUpdateStateFast.f90:71-188);
 *   - block indices are 1-based, as in the reference;
 *   - errors are returned as status codes instead of calling stop_mpi; the
 *     Fortran bridge turns a non-zero code into stop_mpi.
 *
 * Threading: one host thread per handle; device work is ordered on the
 * handle's stream.  All calls are asynchronous except those that return
 * data to the host (download, get_*, set_global_timestep with DtOut).
 */
#ifndef BATS_HIP_H
#define BATS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference's stop_mpi sites they replace) */
enum {
  BATS_OK = 0,
  BATS_EINVAL = 1,      /* bad argument / configuration */
  BATS_ENEGFAST2 = 2,   /* negative fast speed squared,
                           src/ModFaceFlux.f90:4119-4128 */
  BATS_EHLLD = 3,       /* negative HLLD density ratio,
                           src/ModFaceFlux.f90:2985-2997,3067-3078 */
  BATS_ENAN = 4,        /* NaN in the state, src/ModUpdateState.f90:2003-2040 */
  BATS_EBC = 5,         /* unknown boundary condition,
                           src/ModUpdateStateFast.f90:1272 */
  BATS_EHIP = 6,        /* HIP runtime error */
  BATS_ENOMEM = 7,      /* device allocation failed */
  BATS_ENODEV = 8,      /* no usable gfx950 device */
  BATS_ELIMIT = 9       /* request beyond a fixed buffer limit */
};

/* TypeLimiter (src/ModFaceValue.f90:3713-3789) */
enum { BATS_LIM_MINMOD = 1, BATS_LIM_MC = 2, BATS_LIM_MC3 = 3,
       BATS_LIM_BETA = 4, BATS_LIM_NO = 5 };
/* TypeFlux (src/ModFaceFlux.f90:396-404) */
enum { BATS_FLUX_RUSANOV = 1, BATS_FLUX_LINDE = 2, BATS_FLUX_HLLD = 3 };

/* Scheme parameters: the subset of #SCHEME, #TIMESTEPPING, #GAMMA,
 * #MINIMUMPRESSURE, #DBTRICK, #FIXEDTIMESTEP, #DIVB and the compile-time grid
 * shape (Config.pl -g, -ng) that the ideal-MHD explicit update reads
 * (src/ModSetParameters.f90:628-633,1742-1779,2314-2343,3873-3906). */
typedef struct bats_config {
  int nI, nJ, nK;       /* cells per block: 8x8x8 or 16x16x16 */
  int nG;               /* ghost layers of the State_VGB exchange format (2) */
  int nDim;             /* must be 3 for the GPU path */
  int nBlock;           /* blocks held by this handle */
  int nOrder;           /* 1 or 2 */
  int iLimiter;         /* BATS_LIM_* (ignored for nOrder=1) */
  double LimiterBeta;
  int iFlux;            /* BATS_FLUX_* */
  int nStage;           /* 1, 2 or 3 (3: SSP-RK3, needs UseHalfStep = 0,
                           src/ModSetParameters.f90:628-633; the interpolation
                           weights of src/ModUpdateState.f90:748-756) */
  int UseHalfStep;      /* 1: midpoint "half step" scheme (default);
                           0: SSP-RK2 interpolation (#RK) */
  double Cfl;
  double Gamma;
  double pMin;          /* pressure floor, < 0 disables (#MINIMUMPRESSURE) */
  int UseDbTrick;       /* #DBTRICK: add |dB|^2/2 to the energy at stage 1
                           of a two-stage half step scheme and at the one
                           stage of a steady-state run; never for SSP-RK
                           (UseHalfStep = 0) or a one-stage time-accurate run
                           (UseDbTrickNow, src/ModSetParameters.f90:
                           3903-3906; src/ModUpdateState.f90:916-918) */
  int UseDtFixed;       /* #FIXEDTIMESTEP; forces Cfl = 1
                           (src/ModSetParameters.f90:4314-4321) */
  double DtFixed;
  double cTiny;         /* HLLD degeneracy guard (share ModNumConst) */
  int iDevice;          /* HIP device ordinal */
  int UseDivbSource;    /* 8-wave (Powell) div B source (#DIVB,
                           src/ModMain.f90:204; 0 = #DIVB F) */
  int StatePrecision;   /* bytes per state value on the device: 8 (or 0)
                           FP64; 4 = FP32 state and FP32 face arithmetic
                           with the stage update, the pressure fix and the
                           time step in FP64 (no counterpart in the
                           reference, which runs FP64 throughout); the
                           per-block routines need FP64 */
  /* #ADAPTIVELOWORDER (src/ModFaceValue.f90:214-225): at the first stage of
   * each step every face gets a low order weight from the normal velocity
   * jump and the total-pressure ratio of the four cells across it
   * (set_low_order_face, :2699-2965); with nLowOrder = 1 the second order
   * face values use the limited slopes times (1 - weight); nLowOrder = 2
   * leaves the second order scheme unchanged.  pCritLow >= pCritHigh. */
  int UseAdaptiveLowOrder;
  int nLowOrder;        /* 1 or 2 */
  double pCritLow, pCritHigh, VelCrit;
  /* #RESOLUTIONCHANGE (src/ModFaceValue.f90:75-76,168-177): when
   * LimiterBeta > BetaLimiterResChange, the slopes of the cells within
   * nFaceLimiterResChange faces of a block side whose DiLevel_EB is not 0
   * (an outer boundary or a resolution change) use BetaLimiterResChange
   * (:1982-2015).  The reference's defaults are 2.0 and 2;
   * nFaceLimiterResChange = 0 turns it off. */
  double BetaLimiterResChange;
  int nFaceLimiterResChange;
  /* #OUTFLOWPRESSURE: outflow sides (BATS_BC_OUTFLOW) float except for the
   * ghost pressure, which is pOutflow (src/ModCellBoundary.f90:273-277) */
  int UseOutflowPressure;
  double pOutflow;
  /* #TIMEACCURATE F (the reference's IsTimeAccurate = .false.): steady
   * state; every cell advances with Cfl*iStage/nStage times its own
   * DtMax_CB, set at the first stage from that stage's fluxes before the
   * update (src/ModAdvanceExplicit.f90:184-195, src/ModUpdateState.f90:332;
   * the fast path's :562-567); the simulation time does not advance.  The
   * per-block routines follow the reference's order: calc_timestep before
   * update_state at stage 1. */
  int IsSteadyState;
  /* #RESOLUTIONCHANGE UseTvdResChange with first-order prolongation (the
   * reference's default treatment at resolution changes, nOrderProlong =
   * 1): the fine block's ghost layers hold the two coarse cells behind the
   * face, the coarse block's first ghost layer is restricted from the fine
   * cells at the face and moved to the coarse cell distance
   * (correct_monotone_restrict, src/ModFaceValue.f90:233-408), and the
   * fine block's face values at the coarse side and one face in come from
   * tvd_reschange (:680-852,2269-2343).  Two-level grids, fused entry
   * points and per-block routines, with or without the 8-wave source, B0
   * and the inner body, time accurate or steady state; not with low order
   * faces or the surrogate
   * (BATS_EINVAL); 0 keeps this project's second-order prolongation. */
  int UseTvdResChange;
} bats_config;

typedef struct bats_handle bats_handle;

/* Allocate device state for cfg->nBlock blocks.  Replaces init_mod_advance
 * (src/ModAdvance.f90:249-364) for the hot-path arrays. */
int bats_init(const bats_config *cfg, bats_handle **h);
/* Release everything owned by the handle (clean_mod_advance). */
int bats_finalize(bats_handle *h);

/* Grid description for same-level blocks.  iNeighbor_EB[6*b + side]
 * (side = x-,x+,y-,y+,z-,z+) is the 0-based index of the neighbouring
 * block or -1 for an outer boundary: float (zero-gradient,
 * src/ModCellBoundary.f90:518-559) unless bats_set_cell_boundary makes the
 * side fixed; periodic sides point at the wrapped
 * neighbour.  CellSize_DB[3*b + d] = cell size of block b in dimension d.
 * Replaces the BATL tree/geometry queries (DiLevel_EB, CellSize_DB,
 * CellFace_DB, CellVolume_B). */
int bats_set_grid(bats_handle *h, const int *iNeighbor_EB,
                  const double *CellSize_DB);

/* Two-level block-adaptive grid: as bats_set_grid, plus resolution-change
 * sides.  iNeighbor_EB[6*b + side] = -2 - t names slab t of iSlab_IS
 * (8 ints per slab: block, side, kind, n0, n1, n2, n3, target).  Kind 0
 * (the neighbour is coarser): n0 = the coarse block, (n1, n2) = offset of
 * this block's face within the coarse face in coarse cells (0 or nI/2),
 * target = the coarse block's kind-1 slab.  Kind 1 (finer neighbours):
 * n0..n3 = the fine blocks covering the face's quadrants, in-plane
 * coordinates (j,k) / (i,k) / (i,j) on x / y / z sides, quadrant index
 * (p >= nI/2) + 2*(q >= nI/2).  Before every stage the ghost slabs are
 * filled (mean of the 2x2x2 fine cells for coarse ghosts, limited linear
 * prolongation for fine ghosts) and the coarse faces of kind-1 sides take
 * the sum of the four fine face fluxes, as save_cons_flux / apply_cons_flux
 * do (src/ModConserveFlux.f90:79-441, src/ModAdvanceExplicit.f90:95-128).
 * With the 8-wave source, the normal field of the coarse face's left and
 * right states is likewise the mean of the four fine faces' (:355-372),
 * which the div B of the coarse cells next to the face reads.
 * Reallocates the state: call before uploading it. */
int bats_set_amr_grid(bats_handle *h, const int *iNeighbor_EB,
                      const double *CellSize_DB, int nSlab,
                      const int *iSlab_IS);

/* B0 splitting (UseB0, src/ModB0.f90): from here on the state's field is
 * B1 = B - B0 for the static field B0 given at the cell centres and ghost
 * cells, B0_DGB(3, MinI:MaxI, MinJ:MaxJ, MinK:MaxK, nBlock) (Fortran order,
 * nG ghost layers, as the reference's B0_DGB).  The face B0 is the mean of
 * the two cells' B0 (set_b0_face, :314-336); the fluxes, wave speeds and
 * HLLD use the full field where the reference does
 * (src/ModFaceFlux.f90:1203-1212,1798-1805,2905-2958,3111-3131,3912-3920);
 * with UseDivbSource the 8-wave source takes the B0 terms
 * (src/ModCalcSource.f90:1792-1808,1880-1885) and, with UseB0Source,
 * -B1 div(B0) - curl(B0) x B1 from the face B0 (:953-967, set_b0_source
 * :568-590).  The ghost B0 of the outer sides serves the fixed, inflow and
 * vary ghost cells.  On two-level grids a face of a side with finer
 * neighbours takes the mean of the four fine faces' B0 (set_b0_reschange,
 * src/ModB0.f90:338-351,366-455) and the fine blocks' face fluxes that
 * replace the coarse ones carry their B0.  FP64 handles, time accurate or
 * steady state, with the default prolongation at resolution changes or
 * the TVD faces; not with adaptive low order faces (BATS_EINVAL).  A new
 * grid drops B0; B0_DGB = NULL drops it too. */
int bats_set_b0(bats_handle *h, const double *B0_DGB, int UseB0Source);

/* Inner body (UseBody with the 'ionosphere' face boundary condition, as the
 * fast update applies it: src/ModUpdateStateFast.f90:230-265,343-345,
 * 924-937,1277-1327; replaces the body part of set_face_boundary,
 * src/ModFaceBoundary.f90:621-648,876-884).  Used_GB(MinI:MaxI, MinJ:MaxJ,
 * MinK:MaxK, nBlock) (Fortran order, one byte per cell, nonzero: true cell)
 * flags the true cells of every block and its ghost cells (Used_GB of
 * src/ModGeometry.f90); XyzStart_DB(3, nBlock) is the centre of each
 * block's first cell relative to the body centre.  On faces whose limiter
 * stencil meets a body cell the face value becomes first order; a face
 * between a true and a body cell takes the ghost state: density toward
 * BodyRho by at most 10 % of the true value, floating pressure, reversed
 * velocity, B1r times B1rCoef (src/ModFaceBoundary.f90:30, -1 by
 * default); faces between two body cells carry no flux; body cells are not
 * updated and hold DtMax_CB = 0 outside the block minimum
 * (src/ModTimeStepControl.f90:158-159,255).  Works with B0 and the 8-wave
 * source, time accurate or steady state, on same-level and two-level
 * grids (the fine faces that replace coarse ones get the same treatment),
 * through the fused entry points and the per-block routines
 * (bats_calc_face_value applies the face treatment that set_face_boundary
 * applies after calc_face_value in the reference, calc_face_flux zeroes
 * the body-body faces, update_state keeps body cells, calc_timestep masks
 * them), with the default prolongation at resolution changes or the TVD
 * faces; FP64 handles, not with adaptive low order faces or the
 * resolution-change beta (BATS_EINVAL).
 * A new grid drops the body; Used_GB = NULL drops it too. */
int bats_set_body(bats_handle *h, const unsigned char *Used_GB,
                  const double *XyzStart_DB, double BodyRho, double BodyP,
                  double B1rCoef);

/* iTypeCellBc values of the outer boundaries (src/ModCellBoundary.f90:23-41) */
enum { BATS_BC_PERIODIC = 4, BATS_BC_FLOAT = 5, BATS_BC_OUTFLOW = 6,
       BATS_BC_REFLECT = 7, BATS_BC_FIXED = 9, BATS_BC_INFLOW = 10,
       BATS_BC_VARY = 11 };

/* Outer boundary conditions: iTypeCellBc_I[side] for the sides x-, x+, y-,
 * y+, z-, z+ of the domain, applied where iNeighbor_EB = -1.  Float copies
 * the boundary cell into both ghost layers; fixed and inflow set them to
 * CellState_VI[8*side .. 8*side+7] (rho, rho u, B, p), as
 * set_boundary_for_cell does in set_boundary_fast
 * (src/ModUpdateStateFast.f90:1145-1275, CellState_VI of
 * src/ModPhysics.f90:227,679); outflow floats, with the ghost pressure
 * set to bats_config.pOutflow when UseOutflowPressure (:273-277 of
 * src/ModCellBoundary.f90); reflect mirrors the first two cells with the
 * normal momentum and field negated (set_symm_bc on a Cartesian grid,
 * src/ModCellBoundary.f90:308-331,624-666), refilled before every stage
 * and at exchange_messages; a periodic side must have no -1 neighbour.
 * Vary (BATS_BC_VARY) is fixed with the inflow table of
 * bats_set_inflow_table on the x+ side.  With B0 splitting the ghost
 * cells of fixed, inflow and vary sides hold B1 = CellState B - B0 of the
 * ghost cell (:1264-1266).  Other types return BATS_EBC (the reference
 * stops there, :1272).  The default is float on every side.  The types apply to the current grid at
 * once and to every later one; until bats_set_grid the grid is the default
 * one of bats_init (each block its own domain of unit cells, all six sides
 * outer), so a periodic type needs the grid set first.  Fixed sides keep their ghost layers as
 * slabs after the interior cells of the state, so setting them (and any
 * later bats_set_grid / bats_set_amr_grid on such a handle) reallocates the
 * state: call before uploading it.  CellState_VI may be null when no side
 * is fixed, inflow or vary. */
int bats_set_cell_boundary(bats_handle *h, const int *iTypeCellBc_I,
                           const double *CellState_VI);

/* The time-varying inflow (VaryBC_) of the x+ side in time-accurate runs
 * (set_boundary_fast, src/ModUpdateStateFast.f90:1151-1158, with
 * get_solar_wind_point, src/ModSolarwind.f90:591-740, for a satellite on
 * the boundary): before every stage the side's ghost cells take the table's
 * primitive state at the run's simulation time, linear between the points
 * (V_i + w (V_i+1 - V_i)), the end values outside, velocity turned into
 * momentum.  State_VI(8, nPoint) Fortran order (variable fastest),
 * Time_I strictly increasing.  Vary sides other than x+ (or without a
 * table, or in steady state) keep their CellState_VI, as fixed sides do.
 * nPoint = 0 drops the table. */
int bats_set_inflow_table(bats_handle *h, int nPoint, const double *Time_I,
                          const double *State_VI);

/* With UseGraph = 1, bats_advance captures one time step (the global time
 * step and every stage's kernels) into a hipGraph on its first step and
 * replays it for every later step: one graph launch per step instead of
 * 2-6 kernel launches per stage.  No counterpart in the reference (its
 * GPU path launches every kernel); for launch-bound small grids and
 * ensembles.  Recaptured after any grid, run, launch-shape or surrogate
 * change and when TimeSimulationLimit changes; bats_last_stage_kernel_ms
 * then reports the whole step's device time divided by nStage.  Not with
 * the network surrogate (BATS_EINVAL from bats_advance). */
int bats_set_step_graph(bats_handle *h, int UseGraph);

/* set_low_order_face (src/ModFaceValue.f90:2699-2965) for every block from
 * the state of the last bats_exchange_messages (ghost cells included): the
 * weights the per-block bats_calc_face_value applies.  The whole-grid
 * entry points set them themselves at the first stage of each step.
 * Needs UseAdaptiveLowOrder, FP64 state and a same-level grid. */
int bats_set_low_order_face(bats_handle *h);
/* The low order weights of block iBlock (1-based) as LowOrderCrit_XB
 * (nI+1, nJ, nK), LowOrderCrit_YB (nI, nJ+1, nK), LowOrderCrit_ZB
 * (nI, nJ, nK+1), Fortran order (src/ModAdvance.f90 LowOrderCrit_[XYZ]B),
 * as last set by bats_set_low_order_face or a whole-grid step. */
int bats_get_low_order_crit(bats_handle *h, int iBlock, double *CritX,
                            double *CritY, double *CritZ);

/* Host <-> device state in the reference layout State_VGB(nVar, MinI:MaxI,
 * MinJ:MaxJ, MinK:MaxK, nBlock), Fortran order with the variable index
 * fastest (src/ModAdvance.f90:296).  Upload ignores ghost cells; download
 * writes the interior and fills the face ghost slabs the way
 * exchange_messages would (same-level neighbours, float outer sides).
 * Mirrors sync_cpu_gpu (src/ModUpdateStateFast.f90:71-188). */
int bats_upload_state(bats_handle *h, const double *State_VGB);
int bats_download_state(bats_handle *h, double *State_VGB);
/* Same in the device layout (variable-major, interior only):
 * U[((v*nBlock + b)*nK + k)*nJ + j)*nI + i].  No host transposition. */
int bats_upload_state_soa(bats_handle *h, const double *U);
int bats_download_state_soa(bats_handle *h, double *U);

/* ------------------------------------------------------------------------
 * Per-block routines with the reference's names and arguments.  They work
 * on one block at a time through one set of face/flux buffers, like the
 * reference's threadprivate arrays, and exist for per-stage parity dumps
 * and hosts that keep the reference's per-block loop.  Ghost values come
 * from the snapshot taken by bats_exchange_messages.  On a two-level grid
 * the loop is the reference's (src/ModAdvanceExplicit.f90:93-199): for the
 * blocks with a coarser neighbour calc_face_value(iBlock, 1, 1),
 * calc_face_flux(1, iBlock) and save_cons_flux(iBlock); then
 * message_pass_face; then for every block calc_face_value, calc_face_flux,
 * apply_cons_flux, calc_source, update_state (and calc_timestep). */

/* exchange_messages (src/ModMessagePass.f90:27-290): snapshot the current
 * state as the halo source of the next stage. */
int bats_exchange_messages(bats_handle *h);
/* iStage (ModMain) for the per-block routines */
int bats_set_stage(bats_handle *h, int iStage);
/* calc_face_value(iBlock, DoResChangeOnly, DoMonotoneRestrict)
 * (src/ModFaceValue.f90:854).  With DoResChangeOnly only the faces of the
 * block's sides with a coarser neighbour are set (a no-op on a same-level
 * grid).  The ghost cells of resolution changes are the mean of the 2x2x2
 * fine cells or the limited prolongation of the coarse cell (set by
 * bats_exchange_messages); DoMonotoneRestrict is not read. */
int bats_calc_face_value(bats_handle *h, int iBlock, int DoResChangeOnly,
                         int DoMonotoneRestrict);
/* calc_face_flux(DoResChangeOnly, iBlock) -- reference argument order
 * (src/ModFaceFlux.f90:483); DoResChangeOnly as in calc_face_value. */
int bats_calc_face_flux(bats_handle *h, int DoResChangeOnly, int iBlock);
/* save_cons_flux(iBlock) (src/ModConserveFlux.f90:79-187): the face fluxes
 * (and Vdt) of the block's sides with a coarser neighbour, and the normal
 * field of their left and right states times 1/4, stored for the coarse
 * neighbour's face; a no-op for other blocks and on same-level grids. */
int bats_save_cons_flux(bats_handle *h, int iBlock);
/* message_pass_face(nCorrectedFaceValues, CorrectedFlux_V[XYZ]B) as
 * advance_explicit calls it (src/ModAdvanceExplicit.f90:127-128): the
 * stored sub-face records already sit at the coarse side (one device, one
 * stream), so nothing moves; kept for the reference's call sequence. */
int bats_message_pass_face(bats_handle *h);
/* apply_cons_flux(iBlock) (src/ModConserveFlux.f90:260-441): the faces of
 * the block's sides with finer neighbours take the sum of the four fine
 * sub-face fluxes and Vdt, and their face states the sums of the fine
 * normal field (read by calc_source).  Call after calc_face_flux(0,
 * iBlock). */
int bats_apply_cons_flux(bats_handle *h, int iBlock);
/* calc_source(iBlock) (src/ModCalcSource.f90:245-1025): the 8-wave div B
 * source from the last calc_face_value block's face states; zero unless
 * UseDivbSource.  Call it between calc_face_flux and update_state, as
 * advance_explicit does (src/ModAdvanceExplicit.f90:166). */
int bats_calc_source(bats_handle *h, int iBlock);
/* update_state(iBlock) (src/ModUpdateState.f90:83); at iStage==1 copies the
 * block into StateOld first.  With UseDivbSource it adds the source of the
 * preceding bats_calc_source call for the same block (BATS_EINVAL if there
 * was none). */
int bats_update_state(bats_handle *h, int iBlock);
/* calc_timestep(iBlock, IsPartLocal) (src/ModTimeStepControl.f90:111);
 * IsPartLocal must be 0. */
int bats_calc_timestep(bats_handle *h, int iBlock, int IsPartLocal);

/* Face data of the last calc_face_value / calc_face_flux block, in the
 * reference layouts LeftState_VX(nVar,nI+1,nJ,nK), ..._VY(nVar,nI,nJ+1,nK),
 * ..._VZ(nVar,nI,nJ,nK+1) and Flux_VXI(nFaceValue=14, ...). */
int bats_get_face_values(bats_handle *h, double *LeftX, double *RightX,
                         double *LeftY, double *RightY, double *LeftZ,
                         double *RightZ);
int bats_get_face_fluxes(bats_handle *h, double *FluxX, double *FluxY,
                         double *FluxZ);
/* DtMax_CB(nI,nJ,nK,nBlock) and DtMax_B(nBlock) */
int bats_get_dtmax(bats_handle *h, double *DtMax_CB, double *DtMax_B);

/* ------------------------------------------------------------------------
 * Production entry points (all blocks, fused kernels). */

/* set_global_timestep(TimeSimulationLimit)
 * (src/ModTimeStepControl.f90:306-524): increments nStep, reduces DtMax_B
 * on the device, applies the first-step Dt=0 rule and the stop-time clip,
 * and sets Dt = Cfl*DtMin.  If DtOut is non-NULL the call waits and returns
 * the CFL-scaled Dt; otherwise it stays asynchronous. */
int bats_set_global_timestep(bats_handle *h, double TimeSimulationLimit,
                             double *DtOut);
/* advance_explicit(DoCalcTimestep) (src/ModAdvanceExplicit.f90:18-325):
 * all stages of one time step as fused face-value + face-flux + update
 * kernels over all blocks; with DoCalcTimestep the last stage also computes
 * DtMax_B (calc_timestep), and tSimulation advances by Dt/nStage per stage. */
int bats_advance_explicit(bats_handle *h, int DoCalcTimestep);
/* One stage of advance_explicit (iStage = 1..nStage) over all blocks. */
int bats_advance_stage(bats_handle *h, int iStage, int DoCalcTimestep);
/* BATS_advance (src/ModBatsrusMethods.f90:493-759) for nStep steps:
 * set_global_timestep + advance_explicit per step, replayed from a HIP
 * graph after the first step. */
int bats_advance(bats_handle *h, int nStep, double TimeSimulationLimit);
/* Face-speed pass on the current state without an update: fills DtMax_B
 * like the reference's first (Dt = 0) time step. */
int bats_init_timestep(bats_handle *h);

/* Scalars: Dt (CFL-scaled), tSimulation, nStep */
int bats_get_time(bats_handle *h, double *Dt, double *tSimulation,
                  int *nStep);
/* Wait for queued work and return the first error raised on the device */
int bats_synchronize(bats_handle *h);

/* ------------------------------------------------------------------------
 * Ensembles.  No counterpart in the reference, which advances one solution
 * per process: here many independent runs share the handle's flat block
 * list and advance together in every stage launch, each with its own CFL
 * time step (set_global_timestep reduces per run), simulation time and step
 * counter.  Run r owns the blocks iBlockFirst_R[r] .. iBlockFirst_R[r+1]-1
 * (0-based; iBlockFirst_R[0] = 0, iBlockFirst_R[nRun] = nBlock); no
 * neighbour link may cross a run boundary.  Resets every run's time. */
int bats_set_runs(bats_handle *h, int nRun, const int *iBlockFirst_R);
/* Per-run Dt (CFL-scaled), simulation time and step count (nRun each). */
int bats_get_run_times(bats_handle *h, double *Dt_R, double *tSimulation_R,
                       int *nStep_R);
/* Launch shape of the stage kernel: 0 (default) launches one workgroup per
 * group of blocks; n > 0 launches at most n workgroups that walk the flat
 * block work list persistently. */
int bats_set_persistent_grid(bats_handle *h, int nWorkgroup);
/* Cells per thread of the stage kernel and the order of its work items
 * (the reference fixes its loop shape at compile time: Config.pl -g and
 * the OpenACC gang/vector loops of src/ModUpdateStateFast.f90:219-311).
 * CellsPerThread = NB (default), NB/2 or NB/4: each block's columns are
 * split into that many planes per work item, so small grids fill more of
 * the chip.  UseXcdOrder = 1 hands each of the 8 XCDs a contiguous range
 * of the work list, so neighbouring blocks share an L2.  The work order
 * gives the same bits; a split block evaluates the z face at each part
 * boundary in its window prologue, which rounds within an ulp of the
 * unsplit kernel's face there.  Other values return BATS_EINVAL. */
int bats_set_stage_shape(bats_handle *h, int CellsPerThread, int UseXcdOrder);
/* ------------------------------------------------------------------------
 * Learned face-flux surrogate.  No counterpart in the reference.  A small
 * network (16 -> 128 -> 128 -> 64 -> 8, SiLU; FP32 master weights, BF16
 * weights and activations on the matrix cores, FP32 accumulation) maps the
 * left/right primitive face states, rotated to the face normal and scaled
 * by the face's mean density and larger fast speed (16 inputs: the mean of
 * the two scaled states and the unit direction of their difference j), to
 * the correction that takes the face's Rusanov flux to its HLLD (Linde)
 * flux, in the same frame and scaling and divided by |j| (8 values: mass,
 * normal and two tangential momenta, normal and two tangential field,
 * energy).  The
 * parameter vector is flat: W1 (16 x 128, row-major [in][out]), b1, W2
 * (128 x 128), b2, W3 (128 x 64), b3, W4 (64 x 8), b4 -- 27,464 floats. */
typedef struct bats_mlp bats_mlp;
enum { BATS_MLP_NPARAM = 27464, BATS_MLP_NIN = 16, BATS_MLP_NOUT = 8 };
/* bats_mlp_train flags */
enum { BATS_MLP_STANDARDIZE = 1, BATS_MLP_JUMP_WEIGHT = 2,
       BATS_MLP_NO_GRAPH = 4 };
int bats_mlp_create(bats_mlp **m);
int bats_mlp_free(bats_mlp *m);
/* FP32 master weights in (resets the Adam moments) / out; Grad receives
 * the gradient of the last bats_mlp_loss_grad or training step. */
int bats_mlp_set_params(bats_mlp *m, const float *W);
int bats_mlp_get_params(bats_mlp *m, float *W, float *Grad);
/* Y(nRow x 8) = network(X(nRow x 16)), host arrays */
int bats_mlp_forward(bats_mlp *m, const float *X, float *Y, int nRow);
/* bats_mlp_forward without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The rows are staged through a fixed device array of the
 * library (BATS_MLP_NOALLOC_ROWS rows of inputs and outputs, one copy per
 * device, held until the process ends) instead of per-call buffers; nRow
 * above BATS_MLP_NOALLOC_ROWS returns BATS_ELIMIT.  Results are bitwise
 * those of bats_mlp_forward.  All calls on one device share its array:
 * calls that overlap in time on one device (from several host threads or
 * models) race. */
enum { BATS_MLP_NOALLOC_ROWS = 65536 };
int bats_mlp_forward_noalloc(bats_mlp *m, const float *X, float *Y,
                             int nRow);
/* mean-square loss of the network on (X, T) and its gradient (27,464) */
int bats_mlp_loss_grad(bats_mlp *m, const float *X, const float *T, int nRow,
                       double *Loss, float *Grad);
/* training samples held on the device: inputs n x 16, targets n x 8 and
 * the baseline each target corrects (n x 8; zero for bats_mlp_set_data) */
int bats_mlp_set_data(bats_mlp *m, const float *X, const float *T, long n);
int bats_mlp_get_data(bats_mlp *m, long *n, float *X, float *T, float *B);
/* Harvest samples from a solver's current state (FP64 handle): a seeded
 * random fraction of all faces that pass the smoothness gate, each with the
 * FP64 correction from its Rusanov to its HLLD (Linde) flux as target and
 * its Rusanov flux as baseline; appended to the model's samples. */
int bats_mlp_harvest(bats_mlp *m, bats_handle *h, double Fraction,
                     double BetaMin, double JumpTol, unsigned long long Seed,
                     long *nAdded);
/* Adam with cosine learning-rate decay over nEpoch passes of minibatches;
 * HoldFrac of the samples (seeded split) are held out.  Flags: with
 * BATS_MLP_STANDARDIZE the outputs are trained against targets divided by
 * their RMS (the scale is folded back into the last layer afterwards);
 * with BATS_MLP_JUMP_WEIGHT each sample's squared error is weighted by its
 * target weight w (the jump norm of a harvested face) over the mean w, so
 * the loss follows the error of the flux correction w*T itself.  Each
 * training step is captured once as a hipGraph and replayed per minibatch
 * (the learning rate and Adam's bias corrections are read on the device);
 * BATS_MLP_NO_GRAPH launches the same kernels one by one instead, for
 * tools that do not follow graph launches.  Returns the last
 * minibatch loss and the held-out relative L2 errors, overall and per
 * output, with w each sample's target weight (the jump norm |j| for
 * harvested faces, one otherwise): [0..8] of the predicted flux,
 * ||w(Y-T)|| / ||B+wT||, and [9..17] of the correction, ||w(Y-T)|| /
 * ||wT|| (18 values). */
int bats_mlp_train(bats_mlp *m, int nEpoch, int Batch, double Lr0,
                   double HoldFrac, int Flags, unsigned long long Seed,
                   double *TrainLoss, double *HeldOutRelErr_V);

/* Face fluxes of the stage update on an FP32-state handle.  Mode
 * BATS_SUR_OFF: the plain fused stage kernel.  BATS_SUR_NETWORK: faces that
 * pass the smoothness gate take the network's flux, all others the exact
 * HLLD (Linde) flux.  The gate: for every variable and both cells next to
 * the face, the MC limiter leaves the slope unclipped or the two one-sided
 * jumps, scaled like the network inputs, add up to at most JumpTol; and the
 * plasma beta 2p/B^2 of both face states is at least BetaMin.
 * BATS_SUR_RUSANOV: the same gate with the Rusanov flux on the smooth faces
 * (the method without the network).  Every face still carries one flux, so
 * the update stays conservative.  The model must outlive its use.  A
 * surrogate mode returns BATS_EINVAL with the 8-wave div B source
 * (UseDivbSource), adaptive low order faces, a sharper limiter next to
 * resolution changes or steady state: its stage carries none of them. */
enum { BATS_SUR_OFF = 0, BATS_SUR_NETWORK = 1, BATS_SUR_RUSANOV = 2 };
int bats_set_surrogate(bats_handle *h, bats_mlp *m, int Mode, double BetaMin,
                       double JumpTol);
/* faces evaluated and faces given to the network since bats_set_surrogate */
int bats_surrogate_counts(bats_handle *h, long *nFace, long *nNetwork);

/* check_nan (src/ModUpdateState.f90:2003-2040): BATS_ENAN if any state
 * value is NaN or infinite */
int bats_check_nan(bats_handle *h);
/* Conservation integrals: sum over cells of U*V for rho, rho*u (3), B (3)
 * and total energy (8 values), deterministic order (integrate_domain,
 * src/ModWriteLogSatFile.f90:1948-2020). */
int bats_integrate_domain(bats_handle *h, double *Sum_V);

/* Kernel-time measurement: average duration in ms of the fused stage
 * kernel over the last bats_advance call, measured with HIP events on the
 * handle's stream (0 if none). */
int bats_last_stage_kernel_ms(bats_handle *h, double *ms, int *nLaunch);

/* Human-readable text for a status code */
const char *bats_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
