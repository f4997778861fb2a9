! qk_gw_driver: runs the GPU G0W0/evGW drop-ins from QuAcK-format input files.
!
!   qk_gw_driver <working_dir>
!
! Reads, from <working_dir>:
!   input/methods    the "# G0W0 evGW qsGW scGW" line and the "# Rtest ..." line
!                    (src/QuAcK/read_methods.f90:191-199)
!   input/options    the "# GW:" line, 14 fields (src/QuAcK/read_options.f90:240-266)
!   input/hpc_flags  HPC and GPU switches (src/QuAcK/read_hpc_flags.f90:1-39); the GW
!                    drop-ins run only when the GPU switch (line 4) is T, as the
!                    patched dispatch of src/GW/RGW.f90:85,102 and UGW.f90:82,114
!                    does; otherwise the run belongs to QuAcK's CPU routines, which
!                    this driver does not contain, and it stops with code 3
!   input/molecule   "# nAt nEla nElb nCore nRyd" (PyDuck.py:137-145)
!   int/ENuc.dat     nuclear repulsion (PyDuck.py:150-153)
!   int/MO.dat       nOrb, ERHF and eHF(1:nOrb): the RHF reference QuAcK computes
!                    in-process before RGW (src/QuAcK/RQuAcK.f90:252)
!   int/ERI_MO.bin   ERI_MO(nOrb,nOrb,nOrb,nOrb), raw FP64 stream, the array
!                    AOtoMO_ERI_RHF hands to RGW (src/QuAcK/RQuAcK.f90:459,684)
! and for qsGW also:
!   int/nBas.dat, int/ERI.bin (nBas^4, utils/read_2e_integrals.f90:55-63) and
!   int/Ov.dat, Kin.dat, Nuc.dat (utils/read_1e_integrals.f90:38-94)
!   int/X.bin, int/cHF.bin  X and cHF (nBas x nOrb, raw FP64), which QuAcK
!                    computes in-process before RGW
! and writes test/Rtest.dat with QuAcK's dump format when Rtest is set.  The
! driver holds no dipole integrals: they are zero, and so is the dipole that
! qsGW prints.

program qk_gw_driver

  implicit none
  character(len=512) :: wd, line
  logical :: doG0W0, doevGW, doqsGW, doscGW, dotest, dotU, dotG, doRHF, doUHF
  integer :: maxSCF, max_diis, nfreqs, nAt, nEla, nElb, nCore, nRyd, nC, nR, nO, nV, nS, nOrb, ios
  double precision :: thresh, eta, shift, ENuc, ERHF
  character(len=1) :: a1, a2, a3, a4, a5, a6, a7, a8
  logical :: lin, TDA_W, reg, TDA, singlet, triplet, dophBSE, dophBSE2, doppBSE, dBSE, dTDA
  logical :: switch_hpc, use_gpu
  double precision, allocatable :: eHF(:), ERI(:,:,:,:), dip(:,:,:), eGW(:), eU(:,:), SU(:,:), cU(:,:,:)
  double precision, allocatable :: S(:,:), T(:,:), V(:,:), Hc(:,:), X(:,:), cHF(:,:), PHF(:,:), ERI_AO(:,:,:,:)
  double precision, allocatable :: dipAO(:,:,:), ZNuc(:), rNuc(:,:)
  integer :: nBas
  integer :: nOU(2), nCU(2), nVU(2), nRU(2), nSU(2)

  call get_command_argument(1, wd)
  if(len_trim(wd) == 0) wd = '.'

  ! methods
  doG0W0 = .false.; doevGW = .false.; doqsGW = .false.; dotest = .false.; doRHF = .true.; doUHF = .false.
  open(unit=11, file=trim(wd)//'/input/methods', status='old', action='read')
  do
    read(11,'(A)',iostat=ios) line
    if(ios /= 0) exit
    if(index(line, '# RHF UHF') == 1) then
      ! src/QuAcK/read_methods.f90: the HF line selects the restricted or unrestricted branch
      read(11,*) a1, a2
      doRHF = a1 == 'T'; doUHF = a2 == 'T'
    else if(index(line, '# G0W0') == 1) then
      read(11,*) a1, a2, a3, a4
      doG0W0 = a1 == 'T'; doevGW = a2 == 'T'; doqsGW = a3 == 'T'; doscGW = a4 == 'T'
    else if(index(line, '# Rtest') == 1) then
      read(11,*) a1, a2, a3
      dotest = a1 == 'T'; dotU = a2 == 'T'; dotG = a3 == 'T'
    end if
  end do
  close(11)

  ! options (defaults of src/QuAcK/read_options.f90:242-254)
  maxSCF = 64; thresh = 1d-5; max_diis = 1; lin = .false.; eta = 0d0; TDA_W = .false.; reg = .false.
  TDA = .false.; singlet = .true.; triplet = .true.
  dophBSE = .false.; dophBSE2 = .false.; doppBSE = .false.; dBSE = .false.; dTDA = .true.
  open(unit=12, file=trim(wd)//'/input/options', status='old', action='read')
  do
    read(12,'(A)',iostat=ios) line
    if(ios /= 0) exit
    if(index(line, '# GW:') == 1) then
      read(12,*) maxSCF, thresh, max_diis, a1, eta, a2, a3, nfreqs, shift, a4, a5, a6, a7, a8
      lin = a1 == 'T'; TDA_W = a2 == 'T'; reg = a3 == 'T'
    else if(index(line, '# spin:') == 1) then
      ! src/QuAcK/read_options.f90: TDA singlet triplet
      read(12,*) a1, a2, a3
      TDA = a1 == 'T'; singlet = a2 == 'T'; triplet = a3 == 'T'
    else if(index(line, '# BSE:') == 1) then
      ! src/QuAcK/read_options.f90:300-315: phBSE phBSE2 ppBSE dBSE dTDA
      read(12,*) a1, a2, a3, a4, a5
      dophBSE = a1 == 'T'; dophBSE2 = a2 == 'T'; doppBSE = a3 == 'T'; dBSE = a4 == 'T'; dTDA = a5 /= 'F'
    end if
  end do
  close(12)

  ! GPU switch and the dispatch of the patched RGW/UGW: the option the GPU
  ! drop-ins do not carry (ppBSE) stays on the CPU routines
  call qk_read_hpc_flags(wd, switch_hpc, use_gpu)
  if((doG0W0 .or. doevGW .or. doqsGW) .and. .not.(use_gpu .and. .not. doppBSE)) then
    if(.not. use_gpu) then
      write(*,*) 'qk_gw_driver: use_gpu is F in input/hpc_flags: G0W0/evGW belong to QuAcK''s CPU routines'
    else
      write(*,*) 'qk_gw_driver: ppBSE belongs to QuAcK''s CPU routines'
    end if
    write(*,*) '              (RG0W0/evRGW, UG0W0/evUGW), which this driver does not contain'
    error stop 3
  end if

  ! molecule
  open(unit=13, file=trim(wd)//'/input/molecule', status='old', action='read')
  read(13,*)
  read(13,*) nAt, nEla, nElb, nCore, nRyd
  close(13)
  ! core and Rydberg electrons, frozen in pairs (utils/read_molecule.f90:43-53)
  if(mod(nCore,2) /= 0 .or. mod(nRyd,2) /= 0) then
    write(*,*) 'The number of core and Rydberg electrons must be even!'
    error stop 2
  end if
  nC = nCore/2
  nR = nRyd/2
  if(nEla /= nElb .and. .not. doUHF) then
    write(*,*) 'qk_gw_driver: open-shell molecules need the UHF branch'
    error stop 2
  end if
  nO = nEla

  open(unit=14, file=trim(wd)//'/int/ENuc.dat', status='old', action='read')
  read(14,*) ENuc
  close(14)

  open(unit=15, file=trim(wd)//'/int/MO.dat', status='old', action='read')
  read(15,*) nOrb
  read(15,*) ERHF
  allocate(eHF(nOrb))
  read(15,*) eHF
  close(15)

  nV = nOrb - nO
  nS = (nO - nC)*(nV - nR)
  allocate(ERI(nOrb,nOrb,nOrb,nOrb), dip(nOrb,nOrb,3), eGW(nOrb))
  dip = 0d0
  open(unit=16, file=trim(wd)//'/int/ERI_MO.bin', status='old', action='read', form='unformatted', access='stream')
  read(16) ERI
  close(16)

  if(dotest) open(unit=1231597, file=trim(wd)//'/test/Rtest.dat', status='replace')
  if(dotest .and. doUHF) open(unit=1232584, file=trim(wd)//'/test/Utest.dat', status='replace')

  if(doUHF) then
    ! unrestricted branch (src/GW/UGW.f90:82-84,114-116) on the orbitals of int/MO.dat for both
    ! spins, nEla and nElb electrons: the three spin blocks are the same MO integrals
    nOU = [nEla, nElb]; nCU = nC; nRU = nR; nVU = nOrb - nOU; nSU = (nOU - nCU)*(nVU - nRU)
    allocate(eU(nOrb,2), SU(nOrb,nOrb), cU(nOrb,nOrb,2))
    eU(:,1) = eHF; eU(:,2) = eHF; SU = 0d0; cU = 0d0
    ! the "# spin:" line's second and third answers are spin_conserved and spin_flip for
    ! the unrestricted branch (src/QuAcK/read_options.f90:189-197)
    if(doG0W0) call UG0W0_hip(dotest, .false., .false., .false., dophBSE, TDA_W, TDA, dBSE, dTDA, singlet, triplet, &
                              lin, eta, reg, nOrb, nCU, nOU, nVU, nRU, nSU, ENuc, ERHF, SU, ERI, ERI, ERI, dip, dip, &
                              cU, eU)
    if(doevGW) call evUGW_hip(dotest, maxSCF, thresh, max_diis, .false., .false., .false., dophBSE, TDA_W, TDA, &
                              dBSE, dTDA, singlet, triplet, lin, eta, reg, nOrb, nCU, nOU, nVU, nRU, nSU, ENuc, ERHF, &
                              SU, ERI, ERI, ERI, dip, dip, cU, eU)
    deallocate(eU, SU, cU)
  end if
  if(.not. doRHF) then
    doG0W0 = .false.; doevGW = .false.; doqsGW = .false.
  end if

  if(doG0W0) call RG0W0_hip(dotest, .false., .false., .false., dophBSE, dophBSE2, TDA_W, TDA, dBSE, dTDA, &
                            doppBSE, singlet, triplet, lin, eta, reg, .false., nOrb, nOrb, nC, nO, nV, nR, nS, ENuc, &
                            ERHF, ERI, dip, eHF, eGW)
  if(doevGW) call evRGW_hip(dotest, maxSCF, thresh, max_diis, .false., .false., .false., dophBSE, dophBSE2, TDA_W, &
                            TDA, dBSE, dTDA, doppBSE, singlet, triplet, lin, eta, reg, nOrb, nOrb, nC, nO, &
                            nV, nR, nS, ENuc, ERHF, ERI, dip, eHF)

  if(doqsGW) then
    ! qsGW (src/GW/RGW.f90, after G0W0 and evGW): the AO-basis inputs of qsRGW
    open(unit=18, file=trim(wd)//'/int/nBas.dat', status='old', action='read')
    read(18,*) nBas
    close(18)
    allocate(S(nBas,nBas), T(nBas,nBas), V(nBas,nBas), Hc(nBas,nBas), X(nBas,nOrb), cHF(nBas,nOrb), &
             PHF(nBas,nBas), ERI_AO(nBas,nBas,nBas,nBas), dipAO(nBas,nBas,3), ZNuc(1), rNuc(1,3))
    call qk_read_1e(trim(wd)//'/int/Ov.dat', nBas, S)
    call qk_read_1e(trim(wd)//'/int/Kin.dat', nBas, T)
    call qk_read_1e(trim(wd)//'/int/Nuc.dat', nBas, V)
    Hc = T + V
    open(unit=19, file=trim(wd)//'/int/ERI.bin', status='old', action='read', form='unformatted', access='stream')
    read(19) ERI_AO
    close(19)
    open(unit=19, file=trim(wd)//'/int/X.bin', status='old', action='read', form='unformatted', access='stream')
    read(19) X
    close(19)
    open(unit=19, file=trim(wd)//'/int/cHF.bin', status='old', action='read', form='unformatted', access='stream')
    read(19) cHF
    close(19)
    PHF = 2d0*matmul(cHF(:,1:nO), transpose(cHF(:,1:nO)))
    dipAO = 0d0; ZNuc = 0d0; rNuc = 0d0
    call qsRGW_hip(dotest, maxSCF, thresh, max_diis, .false., .false., .false., dophBSE, dophBSE2, TDA_W, TDA, &
                   dBSE, dTDA, doppBSE, singlet, triplet, eta, reg, .false., 0, ZNuc, rNuc, ENuc, nBas, nOrb, nC, &
                   nO, nV, nR, nS, ERHF, S, X, T, V, Hc, ERI_AO, ERI, dipAO, dip, PHF, cHF, eHF)
    deallocate(S, T, V, Hc, X, cHF, PHF, ERI_AO, dipAO, ZNuc, rNuc)
  end if

  if(dotest) close(1231597)
  if(dotest .and. doUHF) close(1232584)

end program

! One symmetric one-electron matrix from PyDuck's "mu nu value" lines
! (utils/read_1e_integrals.f90:43-48).
subroutine qk_read_1e(path, n, M)
  implicit none
  character(len=*), intent(in) :: path
  integer, intent(in) :: n
  double precision, intent(out) :: M(n,n)
  integer :: mu, nu, ios
  double precision :: val
  M = 0d0
  open(unit=20, file=path, status='old', action='read')
  do
    read(20,*,iostat=ios) mu, nu, val
    if(ios /= 0) exit
    M(mu,nu) = val
    M(nu,mu) = val
  end do
  close(20)
end subroutine

! The reader of src/QuAcK/read_hpc_flags.f90:1-39: a comment line, the HPC
! switch, a comment line, the GPU switch; both F when the file is absent.
subroutine qk_read_hpc_flags(wd, switch_hpc, use_gpu)
  implicit none
  character(len=*), intent(in) :: wd
  logical, intent(out) :: switch_hpc, use_gpu
  character(len=1) :: ans
  integer :: ios
  switch_hpc = .false.
  use_gpu = .false.
  open(unit=17, file=trim(wd)//'/input/hpc_flags', status='old', action='read', iostat=ios)
  if(ios /= 0) return
  read(17,*,iostat=ios)
  read(17,*,iostat=ios) ans
  if(ios == 0 .and. ans == 'T') switch_hpc = .true.
  read(17,*,iostat=ios)
  read(17,*,iostat=ios) ans
  if(ios == 0 .and. ans == 'T') use_gpu = .true.
  close(17)
end subroutine

! QuAcK's test-dump format (src/test/dump_test_value.f90:19-20): the
! description on one line, the value as F20.15 on the next.
subroutine dump_test_value(branch, description, val)
  implicit none
  character(len=1), intent(in) :: branch
  character(len=*), intent(in) :: description
  double precision, intent(in) :: val
  if(branch == 'R') then
    write(1231597,*) trim(description)
    write(1231597,'(F20.15)') val
  else if(branch == 'U') then
    write(1232584,*) trim(description)
    write(1232584,'(F20.15)') val
  end if
end subroutine
