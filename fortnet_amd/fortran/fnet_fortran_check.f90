!> Reads a model and a set of structures from a text case file, evaluates them on the GPU
!> through the Fortran host layer (fnet_host), and writes values, outputs and forces -- and
!> optionally one mse gradient and steepest-descent step -- to a text result file.
!>
!>   fnet_fortran_check <case file> <result file> [device]
!>
!> Case file (list-directed, in this order): precision (0 FP64, 1 FP32, 2 FP32 with FP64
!> networks); nG; nG lines "type rcut eta rs kappa lambda xi atomid z1 z2"; 1 and the nG means
!> and nG standard deviations of the z-score, or 0; nExtIn and (if nExtIn > 0) the 1-based
!> indices of the external features the networks read; nSpecies; the atomic number of each
!> network; the activation of each network; nLayers; the layer sizes; nParams; nParams x
!> nSpecies parameters, network after network; nStruct; nExt (external features per atom); per
!> structure "nAtom periodic", the lattice vectors (9 values, vector after vector), then
!> "x y z atomicNumber species ext(1:nExt)" per atom (Bohr, 1-based species); finally 1
!> followed by nGlobal, the global targets (nGlobal x nStruct), the atomic targets
!> ((nOut - nGlobal) x nAtom), the atomic weights (nAtom), the datapoint weights (nStruct) and
!> the learning rate, maximal displacement and tolerance of one steepest-descent step, or 0.
!> Empty arrays have no line.
!>
!> Result file: blocks "NAME n1 n2" followed by n1 x n2 values, one per line, in Fortran order:
!> G (nG, nAtom), E (nOut, nStruct), A (nOut, nAtom), F (3 nOut, nAtom), and with a training
!> step LOSS (1, 1), GRAD (nParams, nSpecies), PARAMS (nParams, nSpecies).
program fnet_fortran_check
  use fnet_host
  implicit none

  type(TFnetModel) :: model
  type(TFnetGeometry), allocatable :: geos(:)
  type(TFnetGpu) :: gpu
  character(len=4096) :: caseFile, resultFile, arg
  integer :: fd, fo, precision, ii, iS, iA, nS, nA, nL, nP, per, hasZ, hasTrain, device, nOut
  integer :: nExtIn, nExt, nGlobal, nAtomAll
  real(dp), allocatable :: vals(:,:), atomic(:,:), sums(:,:), forces(:,:), gTargets(:,:), &
      & aTargets(:,:), grad(:,:), params(:,:), atomicWeights(:), dpWeights(:), extRow(:)
  real(dp) :: loss, lr, maxDisp, tol, xyz(3)
  integer :: zz, sp
  logical :: tConverged

  if (command_argument_count() < 2) then
    write(*, '(a)') 'usage: fnet_fortran_check <case file> <result file> [device]'
    error stop 2
  end if
  call get_command_argument(1, caseFile)
  call get_command_argument(2, resultFile)
  device = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) device
  end if

  open(newunit=fd, file=trim(caseFile), status='old', action='read')
  read(fd, *) precision
  read(fd, *) model%nG
  allocate(model%gType(model%nG), model%atomId(model%nG), model%z1(model%nG), &
      & model%z2(model%nG), model%rCut(model%nG), model%eta(model%nG), model%rs(model%nG), &
      & model%kappa(model%nG), model%lambda(model%nG), model%xi(model%nG))
  do ii = 1, model%nG
    read(fd, *) model%gType(ii), model%rCut(ii), model%eta(ii), model%rs(ii), model%kappa(ii), &
        & model%lambda(ii), model%xi(ii), model%atomId(ii), model%z1(ii), model%z2(ii)
  end do
  read(fd, *) hasZ
  if (hasZ /= 0) then
    allocate(model%zMean(model%nG), model%zStd(model%nG))
    read(fd, *) model%zMean
    read(fd, *) model%zStd
  end if
  read(fd, *) nExtIn
  allocate(model%extIn(nExtIn))
  if (nExtIn > 0) read(fd, *) model%extIn
  read(fd, *) nS
  allocate(model%speciesZ(nS), model%activation(nS))
  read(fd, *) model%speciesZ
  read(fd, *) model%activation
  read(fd, *) nL
  allocate(model%dims(nL))
  read(fd, *) model%dims
  read(fd, *) nP
  allocate(model%params(nP, nS))
  read(fd, *) model%params
  nOut = model%dims(nL)
  read(fd, *) nS
  read(fd, *) nExt
  allocate(geos(nS), extRow(nExt))
  do iS = 1, nS
    read(fd, *) nA, per
    geos(iS)%tPeriodic = per /= 0
    read(fd, *) geos(iS)%latVecs
    allocate(geos(iS)%coords(3, nA), geos(iS)%atomicNumbers(nA), geos(iS)%species(nA))
    if (nExt > 0) allocate(geos(iS)%ext(nExt, nA))
    do iA = 1, nA
      read(fd, *) xyz, zz, sp, extRow
      geos(iS)%coords(:, iA) = xyz
      geos(iS)%atomicNumbers(iA) = zz
      geos(iS)%species(iA) = sp
      if (nExt > 0) geos(iS)%ext(:, iA) = extRow
    end do
  end do
  nAtomAll = sum([(size(geos(iS)%coords, dim=2), iS = 1, nS)])
  read(fd, *) hasTrain
  if (hasTrain /= 0) then
    read(fd, *) nGlobal
    allocate(gTargets(nGlobal, nS), aTargets(nOut - nGlobal, nAtomAll), dpWeights(nS), &
        & atomicWeights(nAtomAll))
    if (nGlobal > 0) read(fd, *) gTargets
    if (nGlobal < nOut) read(fd, *) aTargets
    read(fd, *) atomicWeights
    read(fd, *) dpWeights
    read(fd, *) lr, maxDisp, tol
  end if
  close(fd)

  call gpu%init(device)
  call gpu%uploadModel(model)
  call gpu%setGeometries(geos, precision)
  call gpu%calculate(vals)
  call gpu%predictBatch(atomic, sums)
  call gpu%forceAnalysis(forces)

  open(newunit=fo, file=trim(resultFile), status='replace', action='write')
  call writeBlock(fo, 'G', vals)
  call writeBlock(fo, 'E', sums)
  call writeBlock(fo, 'A', atomic)
  call writeBlock(fo, 'F', forces)
  if (hasTrain /= 0) then
    call gpu%trainGradients(nGlobal, gTargets, aTargets, atomicWeights, dpWeights, grad, loss)
    call gpu%steepestDescent(lr, maxDisp, tol, params, tConverged)
    write(fo, '(a)') 'LOSS 1 1'
    write(fo, '(es26.17e3)') loss
    call writeBlock(fo, 'GRAD', grad)
    call writeBlock(fo, 'PARAMS', params)
  end if
  close(fo)
  call gpu%destroy()

contains

  subroutine writeBlock(fo, name, array)
    integer, intent(in) :: fo
    character(*), intent(in) :: name
    real(dp), intent(in) :: array(:,:)
    integer :: i1, i2

    write(fo, '(a, 1x, i0, 1x, i0)') name, size(array, dim=1), size(array, dim=2)
    do i2 = 1, size(array, dim=2)
      do i1 = 1, size(array, dim=1)
        write(fo, '(es26.17e3)') array(i1, i2)
      end do
    end do
  end subroutine writeBlock

end program fnet_fortran_check
