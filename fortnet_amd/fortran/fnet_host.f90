!> Fortran host types over libfnet_hip.so: the reference's calls on the hot path, on the GPU.
!>
!>  * TFnetGpu%calculate      -- TAcsf%calculate (acsf.F90:540-639): vals(nG, nAtom) of every
!>                               structure, unstandardised, packed structure after structure;
!>  * TFnetGpu%predictBatch   -- TBpnn%predictBatch / iPredict (bpnn.F90:867-900,1001-1058):
!>                               raw atomic outputs (nOut, nAtom) and their per-structure sums;
!>  * TFnetGpu%forceAnalysis  -- forceAnalysis_analytical (forces.F90:317-425) with nJacobian and
!>                               calculatePrime: forces(3 * nOut, nAtom), energy unit per Bohr;
!>  * TFnetGpu%trainGradients -- updateGradients + sysTrain + the normalisation of TBpnn_update
!>                               (bpnn.F90:394-481,610-704,751-767): FP64 gradient (nParams,
!>                               nSpecies) of the mse loss over global targets (the first
!>                               nGlobal outputs summed per structure) and atomic targets (the
!>                               other outputs per atom), and the loss;
!>  * TFnetGpu%steepestDescent -- SteepDesc_next (steepdesc.F90:125-204) on that gradient.
!>
!> Layouts are the reference's: coordinates (3, nAtom) in Bohr, lattice vectors as the columns
!> of latVecs(3, 3), 1-based species (network) indices, parameters (nParams, nSpecies) in the
!> serialised order of network.F90:397-427, external atomic features ext(nExt, nAtom) of each
!> structure (fnetdata.F90:1141-1162) and the 1-based indices of the ones the networks read
!> after the ACSF (TExtFeaturesConfig, netstat.F90:163-211).  A failing call stops the program with the library's
!> message, as the reference's error() does (message.F90:73-84).
module fnet_host
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only : error_unit
  use fnet_hip_c
  implicit none
  private

  public :: dp, TFnetModel, TFnetGeometry, TFnetGpu
  public :: FNET_FP64, FNET_FP32, FNET_FP32_NET64, FNET_BATCHED, FNET_PER_STRUCT

  integer, parameter :: dp = c_double

  !> netstat contents: ACSF table (TGFunctions), z-score and element networks
  type :: TFnetModel
    integer :: nG = 0
    integer, allocatable :: gType(:), atomId(:), z1(:), z2(:)
    real(dp), allocatable :: rCut(:), eta(:), rs(:), kappa(:), lambda(:), xi(:)
    !> z-score means and standard deviations (both unallocated: no standardisation)
    real(dp), allocatable :: zMean(:), zStd(:)
    !> atomic number of each element network, in BPNN net order
    integer, allocatable :: speciesZ(:)
    integer, allocatable :: dims(:), activation(:)
    !> (nParams, nSpecies)
    real(dp), allocatable :: params(:,:)
    !> 1-based indices of the external atomic features fed to the networks after the ACSF
    !> (unallocated or empty: none)
    integer, allocatable :: extIn(:)
  end type TFnetModel

  !> one structure (TGeometry, typegeometry.F90:24-59)
  type :: TFnetGeometry
    real(dp), allocatable :: coords(:,:)
    logical :: tPeriodic = .false.
    real(dp) :: latVecs(3, 3) = 0.0_dp
    integer, allocatable :: atomicNumbers(:)
    !> 1-based network index of each atom
    integer, allocatable :: species(:)
    !> external atomic features (nExt, nAtom); unallocated: none (every structure of a batch
    !> has the same nExt)
    real(dp), allocatable :: ext(:,:)
  end type TFnetGeometry

  !> one device context, its model and one resident batch of structures
  type :: TFnetGpu
    type(c_ptr) :: ctx = c_null_ptr
    type(c_ptr) :: batch = c_null_ptr
    integer :: nG = 0, nOut = 0, nSpecies = 0, nParams = 0
    integer :: nStruct = 0, nAtom = 0
    !> first atom of each structure (1-based) and one past the last
    integer, allocatable :: iStart(:)
  contains
    procedure :: init => TFnetGpu_init
    procedure :: uploadModel => TFnetGpu_uploadModel
    procedure :: setGeometries => TFnetGpu_setGeometries
    procedure :: setPrecision => TFnetGpu_setPrecision
    procedure :: calculate => TFnetGpu_calculate
    procedure :: predictBatch => TFnetGpu_predictBatch
    procedure :: forceAnalysis => TFnetGpu_forceAnalysis
    procedure :: trainGradients => TFnetGpu_trainGradients
    procedure :: steepestDescent => TFnetGpu_steepestDescent
    procedure :: destroy => TFnetGpu_destroy
  end type TFnetGpu

contains

  !> stops with the library's message when status is not FNET_OK
  subroutine check(this, status, where)
    class(TFnetGpu), intent(in) :: this
    integer(c_int), intent(in) :: status
    character(*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: ii
    integer(c_int) :: rc

    if (status == FNET_OK) return
    msg = ''
    if (c_associated(this%ctx)) then
      rc = fnet_last_error(this%ctx, buf, size(buf, kind=c_int))
      if (rc == FNET_OK) then
        do ii = 1, size(buf)
          if (buf(ii) == c_null_char) exit
          msg(ii:ii) = buf(ii)
        end do
      end if
    end if
    write(error_unit, '(a, i0, a)') 'ERROR! ' // where // ' failed (status ', status, '): ' &
        & // trim(msg)
    error stop 1
  end subroutine check


  !> creates the context on a device (0-based HIP device index)
  subroutine TFnetGpu_init(this, device)
    class(TFnetGpu), intent(inout) :: this
    integer, intent(in) :: device
    integer(c_int) :: rc

    rc = fnet_ctx_create(int(device, c_int), this%ctx)
    if (rc /= FNET_OK) then
      write(error_unit, '(a, i0, a, i0, a)') 'ERROR! fnet_ctx_create failed (status ', rc, &
          & '): no usable HIP device ', device, ' for the GPU path'
      error stop 1
    end if
  end subroutine TFnetGpu_init


  !> uploads the netstat contents (TAcsf%fromFile + TBpnn%fromFile)
  subroutine TFnetGpu_uploadModel(this, model)
    class(TFnetGpu), intent(inout) :: this
    type(TFnetModel), intent(in), target :: model
    type(fnet_model_desc) :: md
    integer(c_int), allocatable, target :: gType(:), atomId(:), z1(:), z2(:), spz(:), dims(:), &
        & act(:)
    real(dp), allocatable, target :: rCut(:), eta(:), rs(:), kappa(:), lambda(:), xi(:), &
        & zMean(:), zStd(:), params(:,:)
    integer(c_int), allocatable, target :: extIdx(:)
    integer(c_int), target :: noExt(1)

    this%nG = model%nG
    this%nSpecies = size(model%speciesZ)
    this%nOut = model%dims(size(model%dims))
    this%nParams = size(model%params, dim=1)
    if (fnet_params_per_species(int(size(model%dims), c_int), int(model%dims, c_int)) &
        & /= this%nParams) then
      write(error_unit, '(a)') 'ERROR! network parameters do not match the layer sizes'
      error stop 1
    end if
    gType = model%gType
    atomId = model%atomId
    z1 = model%z1
    z2 = model%z2
    spz = model%speciesZ
    dims = model%dims
    act = model%activation
    rCut = model%rCut
    eta = model%eta
    rs = model%rs
    kappa = model%kappa
    lambda = model%lambda
    xi = model%xi
    params = model%params
    noExt = 0
    md%n_g = int(model%nG, c_int)
    md%g_type = c_loc(gType)
    md%g_rcut = c_loc(rCut)
    md%g_eta = c_loc(eta)
    md%g_rs = c_loc(rs)
    md%g_kappa = c_loc(kappa)
    md%g_lambda = c_loc(lambda)
    md%g_xi = c_loc(xi)
    md%g_atomid = c_loc(atomId)
    md%g_z1 = c_loc(z1)
    md%g_z2 = c_loc(z2)
    if (allocated(model%zMean) .and. allocated(model%zStd)) then
      zMean = model%zMean
      zStd = model%zStd
      md%z_mean = c_loc(zMean)
      md%z_std = c_loc(zStd)
    else
      md%z_mean = c_null_ptr
      md%z_std = c_null_ptr
    end if
    md%n_ext_in = 0
    md%ext_in_idx = c_loc(noExt)
    if (allocated(model%extIn)) then
      if (size(model%extIn) > 0) then
        extIdx = model%extIn
        md%n_ext_in = int(size(extIdx), c_int)
        md%ext_in_idx = c_loc(extIdx)
      end if
    end if
    md%n_species = int(this%nSpecies, c_int)
    md%species_z = c_loc(spz)
    md%n_layers = int(size(model%dims), c_int)
    md%dims = c_loc(dims)
    md%activation = c_loc(act)
    md%params = c_loc(params)
    call check(this, fnet_model_upload(this%ctx, md), 'fnet_model_upload')
  end subroutine TFnetGpu_uploadModel


  !> uploads the structures as one resident batch (replacing the previous one)
  subroutine TFnetGpu_setGeometries(this, geos, precision)
    class(TFnetGpu), intent(inout) :: this
    type(TFnetGeometry), intent(in) :: geos(:)
    integer, intent(in) :: precision
    type(fnet_batch_desc) :: bd
    integer(c_long_long), allocatable, target :: offsets(:)
    real(dp), allocatable, target :: coords(:,:), latVecs(:,:,:)
    integer(c_int), allocatable, target :: periodic(:), species(:), atnum(:)
    real(dp), allocatable, target :: ext(:,:)
    integer :: iS, nA, iA, nExt

    if (c_associated(this%batch)) call fnet_batch_destroy(this%batch)
    this%batch = c_null_ptr
    this%nStruct = size(geos)
    if (allocated(this%iStart)) deallocate(this%iStart)
    allocate(this%iStart(this%nStruct + 1), offsets(this%nStruct + 1))
    this%iStart(1) = 1
    do iS = 1, this%nStruct
      this%iStart(iS + 1) = this%iStart(iS) + size(geos(iS)%coords, dim=2)
    end do
    this%nAtom = this%iStart(this%nStruct + 1) - 1
    offsets(:) = int(this%iStart - 1, c_long_long)
    allocate(coords(3, this%nAtom), latVecs(3, 3, this%nStruct), periodic(this%nStruct), &
        & species(this%nAtom), atnum(this%nAtom))
    nExt = 0
    if (this%nStruct > 0) then
      if (allocated(geos(1)%ext)) nExt = size(geos(1)%ext, dim=1)
    end if
    allocate(ext(nExt, this%nAtom))
    do iS = 1, this%nStruct
      iA = this%iStart(iS)
      nA = size(geos(iS)%coords, dim=2)
      if (nExt == 0) then
        if (allocated(geos(iS)%ext)) then
          if (size(geos(iS)%ext, dim=1) /= 0) call stopWith('structures with and without '&
              & // 'external atomic features in one batch')
        end if
        cycle
      end if
      if (.not. allocated(geos(iS)%ext)) call stopWith('structures with and without external '&
          & // 'atomic features in one batch')
      if (size(geos(iS)%ext, dim=1) /= nExt .or. size(geos(iS)%ext, dim=2) /= nA) &
          & call stopWith('external atomic features of inconsistent shape')
      ext(:, iA:iA + nA - 1) = geos(iS)%ext
    end do
    do iS = 1, this%nStruct
      iA = this%iStart(iS)
      nA = size(geos(iS)%coords, dim=2)
      coords(:, iA:iA + nA - 1) = geos(iS)%coords
      species(iA:iA + nA - 1) = geos(iS)%species - 1
      atnum(iA:iA + nA - 1) = geos(iS)%atomicNumbers
      latVecs(:,:, iS) = geos(iS)%latVecs
      periodic(iS) = merge(1, 0, geos(iS)%tPeriodic)
    end do
    bd%n_struct = int(this%nStruct, c_int)
    bd%offsets = c_loc(offsets)
    bd%coords = c_loc(coords)
    bd%latvecs = c_loc(latVecs)
    bd%periodic = c_loc(periodic)
    bd%species = c_loc(species)
    bd%atnum = c_loc(atnum)
    bd%n_ext = int(nExt, c_int)
    bd%ext = c_null_ptr
    if (nExt > 0) bd%ext = c_loc(ext)
    call check(this, fnet_batch_upload(this%ctx, bd, this%batch), 'fnet_batch_upload')
    call this%setPrecision(precision)
  end subroutine TFnetGpu_setGeometries


  !> arithmetic (FNET_FP64, FNET_FP32, FNET_FP32_NET64) of the batched launches
  subroutine TFnetGpu_setPrecision(this, precision)
    class(TFnetGpu), intent(inout) :: this
    integer, intent(in) :: precision

    call check(this, fnet_batch_set_mode(this%batch, int(precision, c_int), FNET_BATCHED), &
        & 'fnet_batch_set_mode')
  end subroutine TFnetGpu_setPrecision


  !> ACSF values vals(nG, nAtom), structures one after another (TAcsf%calculate)
  subroutine TFnetGpu_calculate(this, vals)
    class(TFnetGpu), intent(inout) :: this
    real(dp), intent(out), allocatable, target :: vals(:,:)

    allocate(vals(this%nG, this%nAtom))
    call check(this, fnet_acsf_forward(this%ctx, this%batch, c_loc(vals)), 'fnet_acsf_forward')
  end subroutine TFnetGpu_calculate


  !> raw atomic outputs (nOut, nAtom) and their sums per structure (nOut, nStruct)
  subroutine TFnetGpu_predictBatch(this, atomic, sums)
    class(TFnetGpu), intent(inout) :: this
    real(dp), intent(out), allocatable, target :: atomic(:,:), sums(:,:)

    allocate(atomic(this%nOut, this%nAtom), sums(this%nOut, this%nStruct))
    call check(this, fnet_energy(this%ctx, this%batch, c_loc(sums), c_loc(atomic)), &
        & 'fnet_energy')
  end subroutine TFnetGpu_predictBatch


  !> analytic forces (3 * nOut, nAtom): -d(output)/d(coordinates) per output
  subroutine TFnetGpu_forceAnalysis(this, forces)
    class(TFnetGpu), intent(inout) :: this
    real(dp), intent(out), allocatable, target :: forces(:,:)

    allocate(forces(3 * this%nOut, this%nAtom))
    call check(this, fnet_forces(this%ctx, this%batch, c_loc(forces)), 'fnet_forces')
  end subroutine TFnetGpu_forceAnalysis


  !> mse gradient of the network parameters, normalised as TBpnn_update, and the loss: the first
  !> nGlobal outputs are summed per structure against globalTargets (nGlobal, nStruct), the
  !> other nOut - nGlobal outputs are compared per atom with atomicTargets (nOut - nGlobal,
  !> nAtom); atomic weights (nAtom) and datapoint weights (nStruct) as in the dataset
  subroutine TFnetGpu_trainGradients(this, nGlobal, globalTargets, atomicTargets, &
      & atomicWeights, dpWeights, grad, loss)
    class(TFnetGpu), intent(inout) :: this
    integer, intent(in) :: nGlobal
    real(dp), intent(in) :: globalTargets(:,:), atomicTargets(:,:), atomicWeights(:), &
        & dpWeights(:)
    real(dp), intent(out), allocatable, target :: grad(:,:)
    real(dp), intent(out), target :: loss
    real(dp), allocatable, target :: gt(:,:), at(:,:), aw(:), dw(:)
    type(c_ptr) :: pGt, pAt

    if (nGlobal < 0 .or. nGlobal > this%nOut) call stopWith('invalid number of global targets')
    if (nGlobal > 0) then
      if (size(globalTargets, dim=1) /= nGlobal .or. size(globalTargets, dim=2) /= this%nStruct) &
          & call stopWith('global targets must be (nGlobal, nStruct)')
    end if
    if (nGlobal < this%nOut) then
      if (size(atomicTargets, dim=1) /= this%nOut - nGlobal &
          & .or. size(atomicTargets, dim=2) /= this%nAtom) &
          & call stopWith('atomic targets must be (nOut - nGlobal, nAtom)')
    end if
    gt = globalTargets
    at = atomicTargets
    aw = atomicWeights
    dw = dpWeights
    pGt = c_null_ptr
    pAt = c_null_ptr
    if (nGlobal > 0) pGt = c_loc(gt)
    if (nGlobal < this%nOut) pAt = c_loc(at)
    allocate(grad(this%nParams, this%nSpecies))
    call check(this, fnet_set_loss(this%ctx, FNET_LOSS_MSE), 'fnet_set_loss')
    call check(this, fnet_train_gradients(this%ctx, this%batch, int(nGlobal, c_int), pGt, pAt, &
        & c_loc(aw), c_loc(dw), c_loc(grad), c_null_ptr, c_loc(loss)), 'fnet_train_gradients')
  end subroutine TFnetGpu_trainGradients


  !> one steepest-descent step on the last gradient; returns the new parameters
  subroutine TFnetGpu_steepestDescent(this, learningRate, maxDisp, tol, params, tConverged)
    class(TFnetGpu), intent(inout) :: this
    real(dp), intent(in) :: learningRate, maxDisp, tol
    real(dp), intent(out), allocatable, target :: params(:,:)
    logical, intent(out) :: tConverged
    integer(c_int) :: conv

    allocate(params(this%nParams, this%nSpecies))
    call check(this, fnet_optimizer_step_sd(this%ctx, learningRate, maxDisp, tol, &
        & c_loc(params), conv), 'fnet_optimizer_step_sd')
    tConverged = conv /= 0
  end subroutine TFnetGpu_steepestDescent


  !> stops with a message, as the reference's error() does
  subroutine stopWith(msg)
    character(*), intent(in) :: msg

    write(error_unit, '(a)') 'ERROR! ' // msg
    error stop 1
  end subroutine stopWith


  subroutine TFnetGpu_destroy(this)
    class(TFnetGpu), intent(inout) :: this

    if (c_associated(this%batch)) call fnet_batch_destroy(this%batch)
    if (c_associated(this%ctx)) call fnet_ctx_destroy(this%ctx)
    this%batch = c_null_ptr
    this%ctx = c_null_ptr
  end subroutine TFnetGpu_destroy

end module fnet_host
