/*
 * bats_oracle.h -- CPU oracle for the explicit block-adaptive ideal-MHD stage
 * update of BATS-R-US (face reconstruction, Rusanov/Linde/HLLD face fluxes,
 * flux-divergence stage update, CFL time step).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * Fortran "orig" CPU path, written so that the GPU product can be checked
 * against it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it; the product never links or calls it.
 *
 * Array layouts follow the reference exactly (Fortran column-major, variable
 * index fastest):
 *   State_VGB(nVar, MinI:MaxI, MinJ:MaxJ, MinK:MaxK, nBlock)
 *       -> State[(((b*nKg + k)*nJg + j)*nIg + i)*nVar + v], ghost offsets nG
 *          in each active dimension (src/ModAdvance.f90:296).
 *   LeftState_VX(nVar, nI+1, nJ, nK) etc.            (src/ModAdvance.f90:178-186)
 *   Flux_VXI(nFaceValue=14, nI+1, nJ, nK)            (src/ModAdvance.f90:46-59,318)
 *   DtMax_CB(nI, nJ, nK, nBlock)                      (src/ModAdvance.f90:124-137)
 *
 * Parity status of each function is stated in oracle/README-style header of
 * bats_oracle.c (pinned vs golden fixture / closed forms / "parity unpinned").
 */
#ifndef BATS_ORACLE_H
#define BATS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* 0-based variable and face-value slots; reference 1-based indices are
 * Rho_=1 .. p_=8, Energy_=9, UnFirst_=10, UnLast_=11, Vdt_=12, BnL_=13, BnR_=14
 * (srcEquation/ModEquationMhd.f90:20-35, src/ModAdvance.f90:46-59). */
enum {
  OB_RHO = 0, OB_RHOUX = 1, OB_RHOUY = 2, OB_RHOUZ = 3,
  OB_BX = 4, OB_BY = 5, OB_BZ = 6, OB_P = 7,
  OB_NVAR = 8, OB_ENERGY = 8, OB_NFLUX = 9,
  OB_UNFIRST = 9, OB_UNLAST = 10, OB_VDT = 11, OB_BNL = 12, OB_BNR = 13,
  OB_NFACEVALUE = 14
};

/* TypeLimiter values (src/ModFaceValue.f90:3713-3789) */
enum { OB_LIM_MINMOD = 1, OB_LIM_MC = 2, OB_LIM_MC3 = 3, OB_LIM_BETA = 4,
       OB_LIM_NO = 5 };
/* TypeFlux values (src/ModFaceFlux.f90:396-404) */
enum { OB_FLUX_RUSANOV = 1, OB_FLUX_LINDE = 2, OB_FLUX_HLLD = 3 };

/* iTypeCellBc values of the outer boundaries (src/ModCellBoundary.f90:23-41);
 * 0 (unset) means float.  Fixed and inflow take the side's CellState_VI
 * state in the ghost cells (src/ModUpdateStateFast.f90:1262-1268); reflect
 * mirrors the block's cells with the normal momentum and field negated
 * (src/ModCellBoundary.f90:308-331,624-666, Cartesian grid); outflow is
 * float with the ghost pressure set to pOutflow when UseOutflowPressure
 * (:273-277). */
enum { OB_BC_PERIODIC = 4, OB_BC_FLOAT = 5, OB_BC_OUTFLOW = 6,
       OB_BC_REFLECT = 7, OB_BC_FIXED = 9, OB_BC_INFLOW = 10,
       OB_BC_VARY = 11 };

/* Error codes (the reference calls stop_mpi at these points) */
enum { OB_OK = 0, OB_ENEGFAST2 = 2, OB_EHLLD = 3, OB_EBC = 5 };

typedef struct {
  int nI, nJ, nK;      /* cells per block */
  int nG;              /* ghost layers (2) */
  int nDim;            /* 1, 2 or 3: faces exist only in active dimensions */
  int nOrder;          /* 1 or 2 */
  int iLimiter;        /* OB_LIM_* */
  double LimiterBeta;
  int iFlux;           /* OB_FLUX_* */
  int nStage;          /* 1 or 2 */
  int UseHalfStep;     /* 1: U = Uold + Source (midpoint); 0: SSP-RK interp */
  double Cfl;
  double Gamma;
  double pMin;         /* < 0: no pressure floor */
  int UseDbTrick;
  int UseDtFixed;
  double DtFixed;
  double cTiny;        /* HLLD degeneracy guard (ModNumConst, absent) */
  int UseDivbSource;   /* 8-wave (Powell) div B source, #DIVB */
  /* #ADAPTIVELOWORDER (src/ModFaceValue.f90:214-225): per-face weight of
   * the low order face value from the velocity jump and total-pressure
   * ratio; with the second order scheme it scales the limited slopes by
   * (1 - weight) when nLowOrder == 1 */
  int UseAdaptiveLowOrder;
  int nLowOrder;       /* 1 or 2 */
  double pCritLow, pCritHigh, VelCrit;
  /* outer boundary type of each side (x-,x+,y-,y+,z-,z+) and the ghost
   * state of the fixed/inflow sides, CellState_VI(nVar, 6) column-major
   * (src/ModPhysics.f90:227,679) */
  int iTypeCellBc_I[6];
  double CellState_VI[6 * OB_NVAR];
  /* #RESOLUTIONCHANGE (src/ModFaceValue.f90:75-76,168-177): when
   * LimiterBeta > BetaLimiterResChange, the slopes of the cells within
   * nFaceLimiterResChange faces of a side with DiLevel_EB /= 0 use
   * BetaLimiterResChange (:1982-2015); reference defaults 2.0 and 2 */
  double BetaLimiterResChange;
  int nFaceLimiterResChange;
  /* #OUTFLOWPRESSURE (src/ModPhysics.f90:731-738 area): outflow sides fix
   * the ghost pressure */
  int UseOutflowPressure;
  double pOutflow;
  /* #TIMEACCURATE F (IsTimeAccurate = .false.): steady state, every cell
   * advances with its own CFL step, calc_timestep runs at the first stage
   * before the update (src/ModAdvanceExplicit.f90:184-195) and
   * set_global_timestep is not called (src/ModBatsrusMethods.f90:569-572) */
  int IsSteadyState;
  /* #RESOLUTIONCHANGE UseTvdResChange with nOrderProlong = 1: two coarse
   * ghost layers on fine blocks, face restriction + correct_monotone_restrict
   * on coarse blocks, tvd_reschange face values on fine blocks
   * (src/ModFaceValue.f90:233-408,680-852,1096-1140,2269-2343) */
  int UseTvdResChange;
} ob_config;

/* --- point routines (exposed for unit tests) ------------------------------ */
/* tvd_reschange (src/ModFaceValue.f90:2269-2343) on primitive states: C2,
 * C1 the two coarse cells behind the face, F1[4][8] / F2[4][8] the fine
 * cells of the first and second layer (subface f = i2 + 2 j2); out: the
 * coarse-side and fine-side values at the four subfaces and the fine
 * cells' values at their next face */
void ob_tvd_reschange(const ob_config *c, const double *C2, const double *C1,
                      const double *F1, const double *F2, double *CtoF,
                      double *FtoC, double *FF);
/* correct_monotone_restrict on one ghost cell: G (primitive, the restricted
 * value at the fine distance) moved to the coarse cell distance from the
 * coarse block's own cell C (src/ModFaceValue.f90:287-294) */
void ob_correct_monotone_restrict(const double *C, double *G);
/* variants with the face's B0 (B0_D[3], NULL: none) for B0 splitting */
void ob_physical_flux_b0(const ob_config *c, int iDim, const double *State_V,
                         const double *B0_D, double *StateCons_V,
                         double *Flux_V, double *UnOut);
int ob_speed_max_b0(const ob_config *c, int iDim, const double *State_V,
                    const double *B0_D, double *Cmax, double *Cleft,
                    double *Cright);
int ob_numerical_flux_b0(const ob_config *c, int iDim, double Area,
                         const double *StateLeft_V,
                         const double *StateRight_V, const double *B0_D,
                         double *Flux_V);
int ob_calc_face_flux_b0(const ob_config *c, const double *CellSize,
                         double *LeftX, double *RightX, double *LeftY,
                         double *RightY, double *LeftZ, double *RightZ,
                         const double *B0X, const double *B0Y,
                         const double *B0Z, double *FluxX, double *FluxY,
                         double *FluxZ);
/* set_b0_face (src/ModB0.f90:314-336): face B0 B0_DX(3, nI+1, nJ, nK) etc.
 * from one block's ghosted B0_DGB(3, ...) */
void ob_set_b0_face(const ob_config *c, const double *B0Blk, double *B0X,
                    double *B0Y, double *B0Z);
/* set_b0_source (src/ModB0.f90:545-590, Cartesian): DivB0_C(nI,nJ,nK),
 * CurlB0_DC(3,nI,nJ,nK) */
void ob_set_b0_source(const ob_config *c, const double *CellSize,
                      const double *B0X, const double *B0Y, const double *B0Z,
                      double *DivB0, double *CurlB0);
/* calc_source with B0: the 8-wave source plus the B0 terms
 * (src/ModCalcSource.f90:1694-1889 with UseB0, :930-967) */
void ob_calc_source_b0(const ob_config *c, const double *CellSize,
                       const double *StateBlk, const double *LeftX,
                       const double *RightX, const double *LeftY,
                       const double *RightY, const double *LeftZ,
                       const double *RightZ, const double *B0Blk,
                       const double *B0X, const double *B0Y,
                       const double *B0Z, const double *DivB0,
                       const double *CurlB0, int UseB0Source,
                       double *Source_VC);
/* limiter on a 1-D pencil: P holds primitives for l = lMin-2 .. lMax+1 with
 * P[(l - lOff)*8 + v]; slopes are written for l = lMin-1 .. lMax into
 * dL/dR[(l - lOff)*8 + v].  Mirrors src/ModFaceValue.f90:3689-3791. */
void ob_limiter(int iLimiter, int lMin, int lMax, double Beta,
                const double *P, int lOff, double *dL, double *dR);

/* physical flux of primitive state (src/ModFaceFlux.f90:1166-1410,1550-1899),
 * iDim = 0,1,2 (Cartesian face normal) */
void ob_physical_flux(const ob_config *c, int iDim, const double *State_V,
                      double *StateCons_V, double *Flux_V, double *Un);
/* speeds of state (src/ModFaceFlux.f90:3545-3722,3853-4252) */
int ob_speed_max(const ob_config *c, int iDim, const double *State_V,
                 double *Cmax, double *Cleft, double *Cright);
/* numerical flux at one face: inputs are the L/R primitive face states,
 * Area; Flux_V has nFaceValue=14 slots (flux*Area in 0..8, Un*Area in 9,
 * Vdt in 11).  src/ModFaceFlux.f90:2130-2473 */
int ob_numerical_flux(const ob_config *c, int iDim, double Area,
                      const double *StateLeftIn_V, const double *StateRightIn_V,
                      double *Flux_V);

/* --- per-block routines --------------------------------------------------- */
/* calc_face_value (src/ModFaceValue.f90:854-2165) for one ghosted block. */
void ob_calc_face_value(const ob_config *c, const double *StateBlk,
                        double *LeftX, double *RightX,
                        double *LeftY, double *RightY,
                        double *LeftZ, double *RightZ);
/* calc_face_value with the low order weights of the block's faces
 * (LowOrderCrit_[XYZ]B; NULL: all zero), applied as
 * Left = P(i-1) + dL(i-1)*(1-w), Right = P(i) - dR(i)*(1-w) when the
 * adaptive low order scheme is on with nLowOrder == 1 and nOrder == 2
 * (src/ModFaceValue.f90:2017-2027,2082-2092,2147-2157) */
void ob_calc_face_value_lo(const ob_config *c, const double *StateBlk,
                           const double *CritX, const double *CritY,
                           const double *CritZ, double *LeftX,
                           double *RightX, double *LeftY, double *RightY,
                           double *LeftZ, double *RightZ);
/* calc_face_value with the low order weights and the sides of the block
 * whose DiLevel_EB is not zero (IsResChange_S[6], x-,x+,y-,y+,z-,z+; NULL:
 * none), where the sharper BetaLimiterResChange applies */
void ob_calc_face_value_rc(const ob_config *c, const double *StateBlk,
                           const double *CritX, const double *CritY,
                           const double *CritZ, const int *IsResChange_S,
                           double *LeftX, double *RightX, double *LeftY,
                           double *RightY, double *LeftZ, double *RightZ);
/* set_phys_based_low_order_face (src/ModFaceValue.f90:2839-2911) for one
 * ghosted block: CritX (nI+1,nJ,nK), CritY (nI,nJ+1,nK), CritZ
 * (nI,nJ,nK+1), each face from the four cells iFace-2 .. iFace+1 along its
 * normal through low_order_face_criteria (:2913-2963) */
void ob_set_low_order_face(const ob_config *c, const double *StateBlk,
                           double *CritX, double *CritY, double *CritZ);
/* the adaptive low order weights change the second order face values */
int ob_low_order_active(const ob_config *c);
/* calc_face_flux (src/ModFaceFlux.f90:483-977); CellSize = dx,dy,dz */
int ob_calc_face_flux(const ob_config *c, const double *CellSize,
                      double *LeftX, double *RightX,
                      double *LeftY, double *RightY,
                      double *LeftZ, double *RightZ,
                      double *FluxX, double *FluxY, double *FluxZ);
/* calc_source (src/ModCalcSource.f90:245-262,862-1025): the Powell div B
 * source on a Cartesian grid without B0.  Source_VC(nFlux=9, nI, nJ, nK),
 * variable fastest; zero unless UseDivbSource. */
void ob_calc_source(const ob_config *c, const double *CellSize,
                    const double *StateBlk, const double *LeftX,
                    const double *RightX, const double *LeftY,
                    const double *RightY, const double *LeftZ,
                    const double *RightZ, double *Source_VC);
/* update_state (src/ModUpdateState.f90:83-965) for one block.  DtMax_CB is
 * this block's (nI,nJ,nK) array; Dt is the CFL-scaled global step;
 * Source_VC from calc_source or NULL (no source terms). */
void ob_update_state(const ob_config *c, int iStage, double Dt,
                     const double *CellSize, const double *DtMax_CB,
                     const double *FluxX, const double *FluxY,
                     const double *FluxZ, const double *Source_VC,
                     double *StateBlk, double *StateOldBlk);
/* calc_timestep (src/ModTimeStepControl.f90:111-304) for one block */
void ob_calc_timestep(const ob_config *c, const double *CellSize,
                      const double *FluxX, const double *FluxY,
                      const double *FluxZ, double *DtMax_CB, double *DtMax_B);

/* --- grid-level routines -------------------------------------------------- */
/* Ghost fill for same-level blocks: iNeighbor_EB[b*6+side] (side order
 * x-,x+,y-,y+,z-,z+) is a 0-based block index or -1 for a float (zero
 * gradient) outer boundary (src/ModCellBoundary.f90:518-559). */
void ob_exchange_messages(const ob_config *c, int nBlock,
                          const int *iNeighbor_EB, double *State);

/* Ghost cells at resolution changes of a two-level grid: restriction to
 * coarse ghosts, limited linear prolongation to fine ghosts (this
 * project's formulas; BATL's are absent).  iSlab_IS: 8 ints per slab
 * [block, side, kind, n0, n1, n2, n3, target] (kind 0: coarser neighbour
 * n0 at in-plane offset (n1, n2) coarse cells; kind 1: finer neighbours
 * n0..n3 by quadrant). */
/* the time-varying inflow state at time t (get_solar_wind_point without a
 * satellite offset, src/ModSolarwind.f90:591-740): the table's primitive
 * states interpolated linearly in time, the end values outside, then the
 * velocity turned into momentum (src/ModUpdateStateFast.f90:1155-1157) */
void ob_vary_state(int nPoint, const double *Time_I, const double *State_VI,
                   double t, double *Cons_V);

void ob_exchange_amr(const ob_config *c, int nBlock, int nSlab,
                     const int *iSlab_IS, double *State);

/* Whole-grid session state for the time loop */
typedef struct {
  ob_config c;
  int nBlock;
  const int *iNeighbor_EB;   /* 6*nBlock */
  const double *CellSize_DB; /* 3*nBlock */
  double *State;             /* ghosted, nBlock blocks */
  double *StateOld;          /* ghosted */
  double *DtMax_CB;          /* nI*nJ*nK*nBlock */
  double *DtMax_B;           /* nBlock */
  double Dt;                 /* CFL-scaled step actually taken */
  double tSimulation;
  int nStep;
  int nThread;               /* OpenMP threads (<=0: default) */
  /* two-level grids: resolution-change slabs (iNeighbor_EB = -2 - slab)
   * and the corrected-flux store of the coarse slabs
   * (nSlab*NB*NB*4*nFaceValue); nSlab = 0 for a same-level grid */
  int nSlab;
  const int *iSlab_IS;
  double *Corr;
  /* low order weights of every block's faces, (nfx + nfy + nfz) per block
   * (X, Y, Z); set at stage 1 of each step when the adaptive low order
   * scheme is on (src/ModAdvanceExplicit.f90:87-91), else unused */
  double *LowOrderCrit;
  /* B0 splitting (UseB0, src/ModB0.f90): the state's field is B1 = B - B0;
   * B0 is B0_DGB(3, MinI:MaxI, MinJ:MaxJ, MinK:MaxK, nBlock), the static
   * field at the cell centres and ghost cells (NULL: no B0); UseB0Source
   * adds -B1 div(B0) - curl(B0) x B1 to the momentum with the 8-wave
   * source (src/ModCalcSource.f90:953-967) */
  const double *B0;
  int UseB0Source;
  /* inner body (UseBody) with the 'ionosphere' face boundary condition:
   * Used_GB of every block (nBlock ghosted blocks of 1 = true cell, 0 =
   * inside the body; NULL: no body), the centre of cell (1,1,1) of each
   * block (XyzStart_DB, 3 per block, body centre at the origin), the body
   * density and pressure (FaceState_VI(:, body1_)) and B1rCoef */
  const unsigned char *Used;
  const double *XyzStart_DB;
  double BodyRho, BodyP, B1rCoef;
  /* time-varying inflow (VaryBC_ on the x+ side, set_boundary_fast,
   * src/ModUpdateStateFast.f90:1151-1158): nVary points of the solar wind
   * table, times VaryTime_I and primitive states VaryState_VI (8 per
   * point); 0: the side keeps its CellState_VI */
  int nVary;
  const double *VaryTime_I, *VaryState_VI;
} ob_session;

/* exchange_messages of a session: same-level and outer ghost cells (with
 * the vary table and B0 at fixed/inflow/vary sides), then the
 * resolution-change ghosts */
void ob_session_exchange(ob_session *s);


/* set_face, 'ionosphere' type (src/ModUpdateStateFast.f90:1277-1327,
 * src/ModFaceBoundary.f90:621-648,876-884): the ghost face state of a face
 * between a true cell and a body cell from the true side's face state;
 * XyzFace_D is the face centre relative to the body centre */
void ob_body_face_state(const double *True_V, const double *XyzFace_D,
                        double BodyRho, double BodyP, double B1rCoef,
                        double *Ghost_V);

/* face states of block b as a stage forms them (with the body's first
 * order fallback and ghost face states); Skip flags body-body faces */
void ob_session_face_values(const ob_session *s, int b, double *LX,
                            double *RX, double *LY, double *RY, double *LZ,
                            double *RZ, unsigned char *SkipX,
                            unsigned char *SkipY, unsigned char *SkipZ);

/* the face B0 (B0_DX(3, nI+1, nJ, nK), B0_DY, B0_DZ), DivB0_C and
 * CurlB0_DC of block b as a stage uses them, with the resolution-change
 * correction on sides with finer neighbours (set_b0_face, set_b0_source);
 * returns -1 without B0 */
int ob_session_b0(const ob_session *s, int b, double *B0X, double *B0Y,
                  double *B0Z, double *DivB0, double *CurlB0);

/* set_global_timestep (src/ModTimeStepControl.f90:306-524) */
void ob_set_global_timestep(ob_session *s, double TimeSimulationLimit);
/* one stage of advance_explicit: all blocks, time advance, ghost exchange
 * between stages (src/ModAdvanceExplicit.f90:74-312) */
int ob_advance_stage(ob_session *s, int iStage, int DoCalcTimestep);
/* every block's 8-wave source from the current state, as a stage computes
 * it (with the resolution-change correction of the face fields) */
int ob_session_sources(ob_session *s, double *Source);
/* advance_explicit (src/ModAdvanceExplicit.f90:18-325), orig CPU path */
int ob_advance_explicit(ob_session *s, int DoCalcTimestep);
/* BATS_advance-shaped time step (src/ModBatsrusMethods.f90:566-653) */
int ob_advance_step(ob_session *s, double TimeSimulationLimit);
/* face-speed pass on the current state only (no update): fills DtMax_CB and
 * DtMax_B exactly as the reference's Dt=0 first step does */
int ob_init_timestep(ob_session *s);

/* pressure <-> energy for one cell (src/ModEnergy.f90:47-127,243-332) */
void ob_pressure_to_energy_cell(const ob_config *c, double *State_V);
void ob_energy_to_pressure_cell(const ob_config *c, double *State_V);

#ifdef __cplusplus
}
#endif
#endif
