!  ModHipBridge -- Fortran host binding of the gfx950 stage-update library
!  (libbats_hip.so, C ABI in include/bats_hip.h).
!
!  The reference's hot path is a set of module procedures
!    calc_face_value(iBlock, DoResChangeOnly, DoMonotoneRestrict)
!                                          src/ModFaceValue.f90:854
!    calc_face_flux(DoResChangeOnly, iBlock) src/ModFaceFlux.f90:483
!    save_cons_flux(iBlock)                src/ModConserveFlux.f90:79
!    message_pass_face                     src/ModAdvanceExplicit.f90:127
!    apply_cons_flux(iBlock)               src/ModConserveFlux.f90:260
!    calc_source(iBlock)                   src/ModCalcSource.f90:245
!    update_state(iBlock)                  src/ModUpdateState.f90:83
!    calc_timestep(iBlock, IsPartLocal)    src/ModTimeStepControl.f90:111
!    set_global_timestep(TimeSimulationLimit)
!                                          src/ModTimeStepControl.f90:306
!    advance_explicit(DoCalcTimestep)      src/ModAdvanceExplicit.f90:18
!    exchange_messages                     src/ModMessagePass.f90:27
!  working on module-global arrays.  This module keeps those names and
!  argument lists; a driver shaped like advance_explicit can use it in place
!  of the reference modules.  The device state lives behind one module-level
!  handle (the analogue of the reference's module globals); State_VGB is
!  exchanged only at sync points (sync_cpu_gpu,
!  src/ModUpdateStateFast.f90:71-188).  A non-zero status from the library
!  stops the program with the library's message, as stop_mpi would.
module ModHipBridge

  use iso_c_binding

  implicit none

  private

  public:: bats_config
  public:: BATS_LIM_MINMOD, BATS_LIM_MC, BATS_LIM_MC3, BATS_LIM_BETA, &
       BATS_LIM_NO, BATS_FLUX_RUSANOV, BATS_FLUX_LINDE, BATS_FLUX_HLLD
  ! raw C entry points
  public:: bats_init, bats_finalize, bats_set_grid, bats_upload_state, &
       bats_download_state, bats_upload_state_soa, bats_download_state_soa, &
       bats_exchange_messages, bats_set_stage, bats_calc_face_value, &
       bats_calc_face_flux, bats_calc_source, bats_update_state, &
       bats_save_cons_flux, bats_message_pass_face, bats_apply_cons_flux, &
       bats_set_b0, bats_set_body, bats_set_inflow_table, &
       bats_calc_timestep, &
       bats_get_face_values, bats_get_face_fluxes, bats_get_dtmax, &
       bats_set_global_timestep, bats_advance_explicit, bats_advance_stage, &
       bats_advance, bats_init_timestep, bats_get_time, bats_synchronize, &
       bats_check_nan, bats_integrate_domain, bats_last_stage_kernel_ms, &
       bats_set_runs, bats_get_run_times, bats_set_persistent_grid, &
       bats_set_stage_shape, &
       bats_set_amr_grid, bats_set_low_order_face, bats_get_low_order_crit, &
       bats_set_step_graph, bats_set_cell_boundary, &
       bats_mlp_create, bats_mlp_free, bats_mlp_set_params, &
       bats_mlp_get_params, bats_mlp_forward, bats_mlp_forward_noalloc, &
       bats_mlp_loss_grad, &
       bats_mlp_set_data, bats_mlp_get_data, bats_mlp_harvest, &
       bats_mlp_train, bats_set_surrogate, bats_surrogate_counts
  ! reference-named routines on the module-level handle
  public:: init_hip_bridge, clean_hip_bridge, sync_state_to_gpu, &
       sync_state_to_cpu, calc_face_value, calc_face_flux, calc_source, &
       save_cons_flux, message_pass_face, apply_cons_flux, &
       update_state, &
       calc_timestep, set_global_timestep, advance_explicit, &
       exchange_messages, set_stage, check_nan, bats_check, bridge_handle, &
       set_low_order_face, set_outer_boundary, &
       init_timestep, get_time
  public:: BATS_BC_PERIODIC, BATS_BC_FLOAT, BATS_BC_OUTFLOW, BATS_BC_REFLECT, &
       BATS_BC_FIXED, BATS_BC_INFLOW

  ! iTypeCellBc values (src/ModCellBoundary.f90:23-41)
  integer(c_int), parameter:: BATS_BC_PERIODIC = 4, BATS_BC_FLOAT = 5, &
       BATS_BC_OUTFLOW = 6, BATS_BC_REFLECT = 7, BATS_BC_FIXED = 9, &
       BATS_BC_INFLOW = 10
  integer(c_int), parameter:: BATS_LIM_MINMOD = 1, BATS_LIM_MC = 2, &
       BATS_LIM_MC3 = 3, BATS_LIM_BETA = 4, BATS_LIM_NO = 5
  integer(c_int), parameter:: BATS_FLUX_RUSANOV = 1, BATS_FLUX_LINDE = 2, &
       BATS_FLUX_HLLD = 3

  ! mirrors struct bats_config of include/bats_hip.h
  type, bind(C):: bats_config
     integer(c_int):: nI, nJ, nK, nG, nDim, nBlock, nOrder, iLimiter
     real(c_double):: LimiterBeta
     integer(c_int):: iFlux, nStage, UseHalfStep
     real(c_double):: Cfl, Gamma, pMin
     integer(c_int):: UseDbTrick, UseDtFixed
     real(c_double):: DtFixed, cTiny
     integer(c_int):: iDevice
     integer(c_int):: UseDivbSource
     integer(c_int):: StatePrecision
     integer(c_int):: UseAdaptiveLowOrder, nLowOrder
     real(c_double):: pCritLow, pCritHigh, VelCrit
     ! #RESOLUTIONCHANGE BetaLimiterResChange, nFaceLimiterResChange
     ! (src/ModFaceValue.f90:75-76; reference defaults 2.0 and 2)
     real(c_double):: BetaLimiterResChange = 2.0_c_double
     integer(c_int):: nFaceLimiterResChange = 2
     ! #OUTFLOWPRESSURE
     integer(c_int):: UseOutflowPressure = 0
     real(c_double):: pOutflow = 0.0_c_double
     ! #TIMEACCURATE F: steady state with per-cell time steps
     integer(c_int):: IsSteadyState = 0
     ! #RESOLUTIONCHANGE UseTvdResChange (first-order prolongation)
     integer(c_int):: UseTvdResChange = 0
  end type bats_config

  interface
     integer(c_int) function bats_init(Config, Handle) &
          bind(C, name="bats_init")
       import:: c_int, c_ptr, bats_config
       type(bats_config), intent(in):: Config
       type(c_ptr), intent(out):: Handle
     end function bats_init
     integer(c_int) function bats_finalize(Handle) &
          bind(C, name="bats_finalize")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_finalize
     integer(c_int) function bats_set_grid(Handle, iNeighbor_EB, CellSize_DB) &
          bind(C, name="bats_set_grid")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       integer(c_int), intent(in):: iNeighbor_EB(*)
       real(c_double), intent(in):: CellSize_DB(*)
     end function bats_set_grid
     integer(c_int) function bats_upload_state(Handle, State_VGB) &
          bind(C, name="bats_upload_state")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(in):: State_VGB(*)
     end function bats_upload_state
     integer(c_int) function bats_download_state(Handle, State_VGB) &
          bind(C, name="bats_download_state")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(inout):: State_VGB(*)
     end function bats_download_state
     integer(c_int) function bats_upload_state_soa(Handle, U) &
          bind(C, name="bats_upload_state_soa")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(in):: U(*)
     end function bats_upload_state_soa
     integer(c_int) function bats_download_state_soa(Handle, U) &
          bind(C, name="bats_download_state_soa")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(inout):: U(*)
     end function bats_download_state_soa
     integer(c_int) function bats_exchange_messages(Handle) &
          bind(C, name="bats_exchange_messages")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_exchange_messages
     integer(c_int) function bats_set_stage(Handle, iStage) &
          bind(C, name="bats_set_stage")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iStage
     end function bats_set_stage
     integer(c_int) function bats_calc_face_value(Handle, iBlock, &
          DoResChangeOnly, DoMonotoneRestrict) &
          bind(C, name="bats_calc_face_value")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock, DoResChangeOnly, DoMonotoneRestrict
     end function bats_calc_face_value
     integer(c_int) function bats_calc_face_flux(Handle, DoResChangeOnly, &
          iBlock) bind(C, name="bats_calc_face_flux")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: DoResChangeOnly, iBlock
     end function bats_calc_face_flux
     integer(c_int) function bats_set_b0(Handle, B0_DGB, UseB0Source) &
          bind(C, name="bats_set_b0")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(in):: B0_DGB(*)
       integer(c_int), value:: UseB0Source
     end function bats_set_b0
     integer(c_int) function bats_set_inflow_table(Handle, nPoint, Time_I, &
          State_VI) bind(C, name="bats_set_inflow_table")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       integer(c_int), value:: nPoint
       real(c_double), intent(in):: Time_I(*), State_VI(*)
     end function bats_set_inflow_table
     integer(c_int) function bats_set_body(Handle, Used_GB, XyzStart_DB, &
          BodyRho, BodyP, B1rCoef) bind(C, name="bats_set_body")
       import:: c_int, c_ptr, c_double, c_signed_char
       type(c_ptr), value:: Handle
       integer(c_signed_char), intent(in):: Used_GB(*)
       real(c_double), intent(in):: XyzStart_DB(*)
       real(c_double), value:: BodyRho, BodyP, B1rCoef
     end function bats_set_body
     integer(c_int) function bats_save_cons_flux(Handle, iBlock) &
          bind(C, name="bats_save_cons_flux")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock
     end function bats_save_cons_flux
     integer(c_int) function bats_message_pass_face(Handle) &
          bind(C, name="bats_message_pass_face")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_message_pass_face
     integer(c_int) function bats_apply_cons_flux(Handle, iBlock) &
          bind(C, name="bats_apply_cons_flux")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock
     end function bats_apply_cons_flux
     integer(c_int) function bats_calc_source(Handle, iBlock) &
          bind(C, name="bats_calc_source")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock
     end function bats_calc_source
     integer(c_int) function bats_update_state(Handle, iBlock) &
          bind(C, name="bats_update_state")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock
     end function bats_update_state
     integer(c_int) function bats_calc_timestep(Handle, iBlock, IsPartLocal) &
          bind(C, name="bats_calc_timestep")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock, IsPartLocal
     end function bats_calc_timestep
     integer(c_int) function bats_get_face_values(Handle, LeftX, RightX, &
          LeftY, RightY, LeftZ, RightZ) bind(C, name="bats_get_face_values")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(inout):: LeftX(*), RightX(*), LeftY(*), &
            RightY(*), LeftZ(*), RightZ(*)
     end function bats_get_face_values
     integer(c_int) function bats_get_face_fluxes(Handle, FluxX, FluxY, &
          FluxZ) bind(C, name="bats_get_face_fluxes")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(inout):: FluxX(*), FluxY(*), FluxZ(*)
     end function bats_get_face_fluxes
     integer(c_int) function bats_get_dtmax(Handle, DtMax_CB, DtMax_B) &
          bind(C, name="bats_get_dtmax")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(inout):: DtMax_CB(*), DtMax_B(*)
     end function bats_get_dtmax
     integer(c_int) function bats_set_global_timestep(Handle, &
          TimeSimulationLimit, DtOut) bind(C, name="bats_set_global_timestep")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), value:: TimeSimulationLimit
       real(c_double), intent(out):: DtOut
     end function bats_set_global_timestep
     integer(c_int) function bats_advance_explicit(Handle, DoCalcTimestep) &
          bind(C, name="bats_advance_explicit")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: DoCalcTimestep
     end function bats_advance_explicit
     integer(c_int) function bats_advance_stage(Handle, iStage, &
          DoCalcTimestep) bind(C, name="bats_advance_stage")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: iStage, DoCalcTimestep
     end function bats_advance_stage
     integer(c_int) function bats_advance(Handle, nStep, TimeSimulationLimit) &
          bind(C, name="bats_advance")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       integer(c_int), value:: nStep
       real(c_double), value:: TimeSimulationLimit
     end function bats_advance
     integer(c_int) function bats_init_timestep(Handle) &
          bind(C, name="bats_init_timestep")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_init_timestep
     integer(c_int) function bats_get_time(Handle, Dt, tSimulation, nStep) &
          bind(C, name="bats_get_time")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(out):: Dt, tSimulation
       integer(c_int), intent(out):: nStep
     end function bats_get_time
     integer(c_int) function bats_synchronize(Handle) &
          bind(C, name="bats_synchronize")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_synchronize
     integer(c_int) function bats_set_runs(Handle, nRun, iBlockFirst_R) &
          bind(C, name="bats_set_runs")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: nRun
       integer(c_int), intent(in):: iBlockFirst_R(*)
     end function bats_set_runs
     integer(c_int) function bats_get_run_times(Handle, Dt_R, &
          tSimulation_R, nStep_R) bind(C, name="bats_get_run_times")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(out):: Dt_R(*), tSimulation_R(*)
       integer(c_int), intent(out):: nStep_R(*)
     end function bats_get_run_times
     integer(c_int) function bats_set_persistent_grid(Handle, nWorkgroup) &
          bind(C, name="bats_set_persistent_grid")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: nWorkgroup
     end function bats_set_persistent_grid
     integer(c_int) function bats_set_stage_shape(Handle, CellsPerThread, &
          UseXcdOrder) bind(C, name="bats_set_stage_shape")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: CellsPerThread, UseXcdOrder
     end function bats_set_stage_shape
     integer(c_int) function bats_set_amr_grid(Handle, iNeighbor_EB, &
          CellSize_DB, nSlab, iSlab_IS) bind(C, name="bats_set_amr_grid")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       integer(c_int), intent(in):: iNeighbor_EB(*), iSlab_IS(*)
       real(c_double), intent(in):: CellSize_DB(*)
       integer(c_int), value:: nSlab
     end function bats_set_amr_grid
     integer(c_int) function bats_set_cell_boundary(Handle, iTypeCellBc_I, &
          CellState_VI) bind(C, name="bats_set_cell_boundary")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       integer(c_int), intent(in):: iTypeCellBc_I(6)
       real(c_double), intent(in):: CellState_VI(8,6)
     end function bats_set_cell_boundary
     integer(c_int) function bats_set_step_graph(Handle, UseGraph) &
          bind(C, name="bats_set_step_graph")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
       integer(c_int), value:: UseGraph
     end function bats_set_step_graph
     integer(c_int) function bats_set_low_order_face(Handle) &
          bind(C, name="bats_set_low_order_face")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_set_low_order_face
     integer(c_int) function bats_get_low_order_crit(Handle, iBlock, &
          LowOrderCrit_X, LowOrderCrit_Y, LowOrderCrit_Z) &
          bind(C, name="bats_get_low_order_crit")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       integer(c_int), value:: iBlock
       real(c_double), intent(out):: LowOrderCrit_X(*), LowOrderCrit_Y(*), &
            LowOrderCrit_Z(*)
     end function bats_get_low_order_crit
     ! learned face-flux surrogate (no counterpart in the reference)
     integer(c_int) function bats_mlp_create(Model) &
          bind(C, name="bats_mlp_create")
       import:: c_int, c_ptr
       type(c_ptr), intent(out):: Model
     end function bats_mlp_create
     integer(c_int) function bats_mlp_free(Model) &
          bind(C, name="bats_mlp_free")
       import:: c_int, c_ptr
       type(c_ptr), value:: Model
     end function bats_mlp_free
     integer(c_int) function bats_mlp_set_params(Model, W) &
          bind(C, name="bats_mlp_set_params")
       import:: c_int, c_ptr, c_float
       type(c_ptr), value:: Model
       real(c_float), intent(in):: W(*)
     end function bats_mlp_set_params
     integer(c_int) function bats_mlp_get_params(Model, W, Grad) &
          bind(C, name="bats_mlp_get_params")
       import:: c_int, c_ptr, c_float
       type(c_ptr), value:: Model
       real(c_float), intent(out):: W(*), Grad(*)
     end function bats_mlp_get_params
     integer(c_int) function bats_mlp_forward(Model, X, Y, nRow) &
          bind(C, name="bats_mlp_forward")
       import:: c_int, c_ptr, c_float
       type(c_ptr), value:: Model
       real(c_float), intent(in):: X(*)
       real(c_float), intent(out):: Y(*)
       integer(c_int), value:: nRow
     end function bats_mlp_forward
     ! the same without device allocation (at most BATS_MLP_NOALLOC_ROWS
     ! rows, staged through the library's fixed per-device array)
     integer(c_int) function bats_mlp_forward_noalloc(Model, X, Y, nRow) &
          bind(C, name="bats_mlp_forward_noalloc")
       import:: c_int, c_ptr, c_float
       type(c_ptr), value:: Model
       real(c_float), intent(in):: X(*)
       real(c_float), intent(out):: Y(*)
       integer(c_int), value:: nRow
     end function bats_mlp_forward_noalloc
     integer(c_int) function bats_mlp_loss_grad(Model, X, T, nRow, Loss, &
          Grad) bind(C, name="bats_mlp_loss_grad")
       import:: c_int, c_ptr, c_float, c_double
       type(c_ptr), value:: Model
       real(c_float), intent(in):: X(*), T(*)
       integer(c_int), value:: nRow
       real(c_double), intent(out):: Loss
       real(c_float), intent(out):: Grad(*)
     end function bats_mlp_loss_grad
     integer(c_int) function bats_mlp_set_data(Model, X, T, n) &
          bind(C, name="bats_mlp_set_data")
       import:: c_int, c_ptr, c_float, c_long
       type(c_ptr), value:: Model
       real(c_float), intent(in):: X(*), T(*)
       integer(c_long), value:: n
     end function bats_mlp_set_data
     integer(c_int) function bats_mlp_get_data(Model, n, X, T, B) &
          bind(C, name="bats_mlp_get_data")
       import:: c_int, c_ptr, c_float, c_long
       type(c_ptr), value:: Model
       integer(c_long), intent(out):: n
       real(c_float), intent(out):: X(*), T(*), B(*)
     end function bats_mlp_get_data
     integer(c_int) function bats_mlp_harvest(Model, Handle, Fraction, &
          BetaMin, JumpTol, Seed, nAdded) bind(C, name="bats_mlp_harvest")
       import:: c_int, c_ptr, c_double, c_long, c_long_long
       type(c_ptr), value:: Model, Handle
       real(c_double), value:: Fraction, BetaMin, JumpTol
       integer(c_long_long), value:: Seed
       integer(c_long), intent(out):: nAdded
     end function bats_mlp_harvest
     integer(c_int) function bats_mlp_train(Model, nEpoch, Batch, Lr0, &
          HoldFrac, Flags, Seed, TrainLoss, HeldOutRelErr_V) &
          bind(C, name="bats_mlp_train")
       import:: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value:: Model
       integer(c_int), value:: nEpoch, Batch, Flags
       real(c_double), value:: Lr0, HoldFrac
       integer(c_long_long), value:: Seed
       real(c_double), intent(out):: TrainLoss, HeldOutRelErr_V(18)
     end function bats_mlp_train
     integer(c_int) function bats_set_surrogate(Handle, Model, Mode, &
          BetaMin, JumpTol) bind(C, name="bats_set_surrogate")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle, Model
       integer(c_int), value:: Mode
       real(c_double), value:: BetaMin, JumpTol
     end function bats_set_surrogate
     integer(c_int) function bats_surrogate_counts(Handle, nFace, &
          nNetwork) bind(C, name="bats_surrogate_counts")
       import:: c_int, c_ptr, c_long
       type(c_ptr), value:: Handle
       integer(c_long), intent(out):: nFace, nNetwork
     end function bats_surrogate_counts
     integer(c_int) function bats_check_nan(Handle) &
          bind(C, name="bats_check_nan")
       import:: c_int, c_ptr
       type(c_ptr), value:: Handle
     end function bats_check_nan
     integer(c_int) function bats_integrate_domain(Handle, Sum_V) &
          bind(C, name="bats_integrate_domain")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(out):: Sum_V(8)
     end function bats_integrate_domain
     integer(c_int) function bats_last_stage_kernel_ms(Handle, Ms, nLaunch) &
          bind(C, name="bats_last_stage_kernel_ms")
       import:: c_int, c_ptr, c_double
       type(c_ptr), value:: Handle
       real(c_double), intent(out):: Ms
       integer(c_int), intent(out):: nLaunch
     end function bats_last_stage_kernel_ms
     type(c_ptr) function bats_strerror_c(Code) bind(C, name="bats_strerror")
       import:: c_int, c_ptr
       integer(c_int), value:: Code
     end function bats_strerror_c
  end interface

  ! the module-level handle, the analogue of the reference's module arrays
  type(c_ptr), save:: hBats = c_null_ptr

contains
  !============================================================================
  subroutine bats_check(iError, NameWhere)
    ! Stop with the library's message on a non-zero status (stop_mpi)
    integer(c_int), intent(in):: iError
    character(len=*), intent(in):: NameWhere
    character(len=200):: String
    character(kind=c_char), pointer:: Msg_I(:)
    integer:: i
    !--------------------------------------------------------------------------
    if(iError == 0) RETURN
    call c_f_pointer(bats_strerror_c(iError), Msg_I, [200])
    String = ' '
    do i = 1, 200
       if(Msg_I(i) == c_null_char) EXIT
       String(i:i) = Msg_I(i)
    end do
    write(*,'(a,i3,2a)') 'ERROR in '//NameWhere//': status ', iError, &
         ': ', trim(String)
    error stop 'stop_mpi'
  end subroutine bats_check
  !============================================================================
  subroutine init_hip_bridge(Config, iNeighbor_EB, CellSize_DB)
    type(bats_config), intent(in):: Config
    integer(c_int), intent(in):: iNeighbor_EB(6,Config%nBlock)
    real(c_double), intent(in):: CellSize_DB(3,Config%nBlock)
    !--------------------------------------------------------------------------
    call bats_check(bats_init(Config, hBats), 'init_hip_bridge')
    call bats_check(bats_set_grid(hBats, iNeighbor_EB, CellSize_DB), &
         'init_hip_bridge')
  end subroutine init_hip_bridge
  !============================================================================
  subroutine clean_hip_bridge
    !--------------------------------------------------------------------------
    if(c_associated(hBats)) call bats_check(bats_finalize(hBats), &
         'clean_hip_bridge')
    hBats = c_null_ptr
  end subroutine clean_hip_bridge
  !============================================================================
  subroutine sync_state_to_gpu(State_VGB)
    real(c_double), intent(in):: State_VGB(:,:,:,:,:)
    !--------------------------------------------------------------------------
    call bats_check(bats_upload_state(hBats, State_VGB), 'sync_state_to_gpu')
  end subroutine sync_state_to_gpu
  !============================================================================
  subroutine sync_state_to_cpu(State_VGB)
    real(c_double), intent(inout):: State_VGB(:,:,:,:,:)
    !--------------------------------------------------------------------------
    call bats_check(bats_download_state(hBats, State_VGB), &
         'sync_state_to_cpu')
  end subroutine sync_state_to_cpu
  !============================================================================
  subroutine exchange_messages
    !--------------------------------------------------------------------------
    call bats_check(bats_exchange_messages(hBats), 'exchange_messages')
  end subroutine exchange_messages
  !============================================================================
  subroutine set_stage(iStage)
    integer, intent(in):: iStage
    !--------------------------------------------------------------------------
    call bats_check(bats_set_stage(hBats, iStage), 'set_stage')
  end subroutine set_stage
  !============================================================================
  subroutine calc_face_value(iBlock, DoResChangeOnly, DoMonotoneRestrict)
    integer, intent(in):: iBlock
    logical, intent(in):: DoResChangeOnly, DoMonotoneRestrict
    !--------------------------------------------------------------------------
    call bats_check(bats_calc_face_value(hBats, iBlock, &
         merge(1, 0, DoResChangeOnly), merge(1, 0, DoMonotoneRestrict)), &
         'calc_face_value')
  end subroutine calc_face_value
  !============================================================================
  subroutine set_outer_boundary(iTypeCellBc_I, CellState_VI)
    ! outer boundary types (#OUTERBOUNDARY) and the ghost states of the
    ! fixed/inflow sides, CellState_VI of src/ModPhysics.f90:227,679;
    ! reallocates the state, so call before sync_state_to_gpu
    integer(c_int), intent(in):: iTypeCellBc_I(6)
    real(c_double), intent(in):: CellState_VI(8,6)
    !--------------------------------------------------------------------------
    call bats_check(bats_set_cell_boundary(hBats, iTypeCellBc_I, &
         CellState_VI), 'set_outer_boundary')
  end subroutine set_outer_boundary
  !============================================================================
  subroutine set_low_order_face
    ! low order face weights of all blocks (src/ModFaceValue.f90:2699)
    !--------------------------------------------------------------------------
    call bats_check(bats_set_low_order_face(hBats), 'set_low_order_face')
  end subroutine set_low_order_face
  !============================================================================
  subroutine calc_face_flux(DoResChangeOnly, iBlock)
    logical, intent(in):: DoResChangeOnly
    integer, intent(in):: iBlock
    !--------------------------------------------------------------------------
    call bats_check(bats_calc_face_flux(hBats, merge(1, 0, DoResChangeOnly), &
         iBlock), 'calc_face_flux')
  end subroutine calc_face_flux
  !============================================================================
  subroutine save_cons_flux(iBlock)
    ! fine-side face fluxes at resolution changes (ModConserveFlux)
    integer, intent(in):: iBlock
    !--------------------------------------------------------------------------
    call bats_check(bats_save_cons_flux(hBats, iBlock), 'save_cons_flux')
  end subroutine save_cons_flux
  !============================================================================
  subroutine message_pass_face
    ! the stored fine-face records already sit at the coarse sides
    !--------------------------------------------------------------------------
    call bats_check(bats_message_pass_face(hBats), 'message_pass_face')
  end subroutine message_pass_face
  !============================================================================
  subroutine apply_cons_flux(iBlock)
    ! coarse-side faces take the sums of the fine face fluxes
    integer, intent(in):: iBlock
    !--------------------------------------------------------------------------
    call bats_check(bats_apply_cons_flux(hBats, iBlock), 'apply_cons_flux')
  end subroutine apply_cons_flux
  !============================================================================
  subroutine calc_source(iBlock)
    integer, intent(in):: iBlock
    !--------------------------------------------------------------------------
    call bats_check(bats_calc_source(hBats, iBlock), 'calc_source')
  end subroutine calc_source
  !============================================================================
  subroutine update_state(iBlock)
    integer, intent(in):: iBlock
    !--------------------------------------------------------------------------
    call bats_check(bats_update_state(hBats, iBlock), 'update_state')
  end subroutine update_state
  !============================================================================
  subroutine calc_timestep(iBlock, IsPartLocal)
    integer, intent(in):: iBlock
    logical, optional, intent(in):: IsPartLocal
    integer(c_int):: iPart
    !--------------------------------------------------------------------------
    iPart = 0
    if(present(IsPartLocal)) iPart = merge(1, 0, IsPartLocal)
    call bats_check(bats_calc_timestep(hBats, iBlock, iPart), 'calc_timestep')
  end subroutine calc_timestep
  !============================================================================
  subroutine set_global_timestep(TimeSimulationLimit, Dt)
    real(c_double), intent(in):: TimeSimulationLimit
    real(c_double), intent(out):: Dt
    !--------------------------------------------------------------------------
    call bats_check(bats_set_global_timestep(hBats, TimeSimulationLimit, Dt), &
         'set_global_timestep')
  end subroutine set_global_timestep
  !============================================================================
  subroutine advance_explicit(DoCalcTimestep)
    logical, intent(in):: DoCalcTimestep
    !--------------------------------------------------------------------------
    call bats_check(bats_advance_explicit(hBats, merge(1, 0, DoCalcTimestep)),&
         'advance_explicit')
  end subroutine advance_explicit
  !============================================================================
  subroutine check_nan
    !--------------------------------------------------------------------------
    call bats_check(bats_check_nan(hBats), 'check_nan')
  end subroutine check_nan
  !============================================================================
  subroutine init_timestep
    ! time step limit from the current state without an update
    !--------------------------------------------------------------------------
    call bats_check(bats_init_timestep(hBats), 'init_timestep')
  end subroutine init_timestep
  !============================================================================
  subroutine get_time(Dt, tSimulation, nStep)
    real(c_double), intent(out):: Dt, tSimulation
    integer(c_int), intent(out):: nStep
    !--------------------------------------------------------------------------
    call bats_check(bats_get_time(hBats, Dt, tSimulation, nStep), 'get_time')
  end subroutine get_time
  !============================================================================
  type(c_ptr) function bridge_handle()
    ! the module-level handle, for callers of the raw bats_* entry points
    bridge_handle = hBats
  end function bridge_handle
  !============================================================================
end module ModHipBridge
!==============================================================================
