!  bats_fortran_driver -- a Fortran host that runs the explicit MHD update on
!  the GPU through ModHipBridge, with the time loop shaped like the
!  reference's BATS_advance / advance_explicit
!  (src/ModBatsrusMethods.f90:566-653, src/ModAdvanceExplicit.f90:74-312).
!
!  Usage: bats_fortran_driver <run directory>
!  The run directory holds
!    PARAM.nml      namelist &RUN (grid shape, scheme, nStep, TypeLoop)
!    grid.bin       iNeighbor_EB(6,nBlock) int32, CellSize_DB(3,nBlock) real8
!    state_in.bin   State_VGB(8,-1:nI+2,-1:nJ+2,-1:nK+2,nBlock) real8
!  and, with UseB0 / UseBody,
!    b0.bin         B0_DGB(3,-1:nI+2,-1:nJ+2,-1:nK+2,nBlock) real8
!    body.bin       Used_GB(-1:nI+2,-1:nJ+2,-1:nK+2,nBlock) int8 (1: true
!                   cell), then XyzStart_DB(3,nBlock) real8
!  and receives state_out.bin (same layout) and log.txt.
!  TypeLoop='blocks' calls the per-block routines in the reference's order;
!  TypeLoop='fused' calls set_global_timestep + advance_explicit.
program bats_fortran_driver

  use iso_c_binding
  use ModHipBridge

  implicit none

  integer:: nRootI = 2, nRootJ = 2, nRootK = 2, nCell = 8, nStep = 2
  integer:: nOrder = 2, nStage = 2, iLimiter = 2, iFlux = 1
  real(c_double):: LimiterBeta = 1.5, Cfl = 0.8, Gamma = 5.0/3.0
  character(len=10):: TypeLoop = 'blocks'
  logical:: UseDivbSource = .false.
  logical:: UseAdaptiveLowOrder = .false.
  integer:: nLowOrder = 1
  real(c_double):: pCritLow = 2.0, pCritHigh = 1.5, VelCrit = 2.0
  ! outer boundaries of the sides where iNeighbor_EB = -1 (float by default)
  integer(c_int):: iTypeCellBc(6) = 5
  real(c_double):: CellState_VI(8,6) = 0.0
  ! #RESOLUTIONCHANGE limiter (reference defaults)
  real(c_double):: BetaLimiterResChange = 2.0
  integer:: nFaceLimiterResChange = 2
  ! #OUTFLOWPRESSURE for outflow sides
  logical:: UseOutflowPressure = .false.
  real(c_double):: pOutflow = 0.0
  ! #TIMEACCURATE (false: steady state, per-cell time steps)
  logical:: IsTimeAccurate = .true.
  ! B0 splitting (B0_DGB from b0.bin) and the inner body (Used_GB and
  ! XyzStart_DB from body.bin) with its face state and B1rCoef
  logical:: UseB0 = .false., UseB0Source = .true., UseBody = .false.
  real(c_double):: BodyRho = 1.0, BodyP = 1.0, B1rCoef = -1.0
  namelist /RUN/ nRootI, nRootJ, nRootK, nCell, nStep, nOrder, nStage, &
       iLimiter, iFlux, LimiterBeta, Cfl, Gamma, TypeLoop, UseDivbSource, &
       UseAdaptiveLowOrder, nLowOrder, pCritLow, pCritHigh, VelCrit, &
       iTypeCellBc, CellState_VI, BetaLimiterResChange, nFaceLimiterResChange, &
       UseOutflowPressure, pOutflow, IsTimeAccurate, UseB0, UseB0Source, &
       UseBody, BodyRho, BodyP, B1rCoef

  type(bats_config):: Config
  character(len=512):: NameDir
  integer:: nBlock, iBlock, iStep, iStage, iUnit, nG
  integer(c_int), allocatable:: iNeighbor_EB(:,:)
  real(c_double), allocatable:: CellSize_DB(:,:), State_VGB(:,:,:,:,:)
  real(c_double), allocatable:: B0_DGB(:,:,:,:,:), XyzStart_DB(:,:)
  integer(c_signed_char), allocatable:: Used_GB(:,:,:,:)
  real(c_double):: Dt, tSimulation
  integer(c_int):: nStepDone
  !----------------------------------------------------------------------------
  call get_command_argument(1, NameDir)
  open(newunit=iUnit, file=trim(NameDir)//'/PARAM.nml', status='old')
  read(iUnit, nml=RUN)
  close(iUnit)

  nG = 2
  nBlock = nRootI*nRootJ*nRootK
  allocate(iNeighbor_EB(6,nBlock), CellSize_DB(3,nBlock))
  allocate(State_VGB(8,1-nG:nCell+nG,1-nG:nCell+nG,1-nG:nCell+nG,nBlock))
  open(newunit=iUnit, file=trim(NameDir)//'/grid.bin', access='stream', &
       form='unformatted', status='old')
  read(iUnit) iNeighbor_EB, CellSize_DB
  close(iUnit)
  open(newunit=iUnit, file=trim(NameDir)//'/state_in.bin', access='stream', &
       form='unformatted', status='old')
  read(iUnit) State_VGB
  close(iUnit)

  Config%nI = nCell; Config%nJ = nCell; Config%nK = nCell
  Config%nG = nG; Config%nDim = 3; Config%nBlock = nBlock
  Config%nOrder = nOrder; Config%iLimiter = iLimiter
  Config%LimiterBeta = LimiterBeta; Config%iFlux = iFlux
  Config%nStage = nStage; Config%UseHalfStep = 1
  Config%Cfl = Cfl; Config%Gamma = Gamma; Config%pMin = -1.0
  Config%UseDbTrick = 0; Config%UseDtFixed = 0; Config%DtFixed = 0.0
  Config%cTiny = 1e-6; Config%iDevice = 0
  Config%UseDivbSource = merge(1, 0, UseDivbSource)
  Config%StatePrecision = 8
  Config%UseAdaptiveLowOrder = merge(1, 0, UseAdaptiveLowOrder)
  Config%nLowOrder = nLowOrder
  Config%pCritLow = pCritLow; Config%pCritHigh = pCritHigh
  Config%VelCrit = VelCrit
  Config%BetaLimiterResChange = BetaLimiterResChange
  Config%nFaceLimiterResChange = nFaceLimiterResChange
  Config%UseOutflowPressure = merge(1, 0, UseOutflowPressure)
  Config%pOutflow = pOutflow
  Config%IsSteadyState = merge(0, 1, IsTimeAccurate)

  call init_hip_bridge(Config, iNeighbor_EB, CellSize_DB)
  if(any(iTypeCellBc /= BATS_BC_FLOAT)) &
       call set_outer_boundary(iTypeCellBc, CellState_VI)
  ! B0 and the body after the grid and boundaries (a new grid drops them)
  if(UseB0)then
     allocate(B0_DGB(3,1-nG:nCell+nG,1-nG:nCell+nG,1-nG:nCell+nG,nBlock))
     open(newunit=iUnit, file=trim(NameDir)//'/b0.bin', access='stream', &
          form='unformatted', status='old')
     read(iUnit) B0_DGB
     close(iUnit)
     call bats_check(bats_set_b0(bridge_handle(), B0_DGB, &
          merge(1, 0, UseB0Source)), 'set_b0')
  end if
  if(UseBody)then
     allocate(Used_GB(1-nG:nCell+nG,1-nG:nCell+nG,1-nG:nCell+nG,nBlock), &
          XyzStart_DB(3,nBlock))
     open(newunit=iUnit, file=trim(NameDir)//'/body.bin', access='stream', &
          form='unformatted', status='old')
     read(iUnit) Used_GB, XyzStart_DB
     close(iUnit)
     ! the body's face treatment is part of calc_face_value here, so the
     ! loop below has no set_face_boundary call
     call bats_check(bats_set_body(bridge_handle(), Used_GB, XyzStart_DB, &
          BodyRho, BodyP, B1rCoef), 'set_body')
  end if
  call sync_state_to_gpu(State_VGB)

  if(TypeLoop == 'blocks')then
     ! time step limit from the initial state (the reference's Dt=0 step)
     call exchange_messages
     if(UseAdaptiveLowOrder) call set_low_order_face
     call set_stage(nStage)
     do iBlock = 1, nBlock
        call calc_face_value(iBlock, .false., .true.)
        call calc_face_flux(.false., iBlock)
        call calc_timestep(iBlock)
     end do
     do iStep = 1, nStep
        call set_global_timestep(-1.0_c_double, Dt)
        STAGELOOP: do iStage = 1, nStage
           call set_stage(iStage)
           ! src/ModAdvanceExplicit.f90:87-91
           if(iStage == 1 .and. UseAdaptiveLowOrder) call set_low_order_face
           do iBlock = 1, nBlock
              call calc_face_value(iBlock, .false., .true.)
              call calc_face_flux(.false., iBlock)
              call calc_source(iBlock)
              ! src/ModAdvanceExplicit.f90:184-195 and :217-219
              if(.not.IsTimeAccurate .and. iStage == 1) &
                   call calc_timestep(iBlock)
              call update_state(iBlock)
              if(IsTimeAccurate .and. iStage == nStage) &
                   call calc_timestep(iBlock)
           end do
           call exchange_messages
        end do STAGELOOP
     end do
  else
     call init_timestep
     do iStep = 1, nStep
        call set_global_timestep(-1.0_c_double, Dt)
        call advance_explicit(.true.)
     end do
  end if
  call check_nan
  call sync_state_to_cpu(State_VGB)
  call get_time(Dt, tSimulation, nStepDone)

  open(newunit=iUnit, file=trim(NameDir)//'/state_out.bin', &
       access='stream', form='unformatted', status='replace')
  write(iUnit) State_VGB
  close(iUnit)
  open(newunit=iUnit, file=trim(NameDir)//'/log.txt', status='replace')
  write(iUnit,'(a,i6,2es24.16)') 'nStep Dt tSimulation ', nStepDone, Dt, &
       tSimulation
  close(iUnit)
  call clean_hip_bridge

end program bats_fortran_driver
!==============================================================================
