! Drop-in GPU forms of QuAcK's restricted G0W0 and evGW drivers.
!
! RG0W0_hip and evRGW_hip take exactly the argument lists of RG0W0
! (src/GW/RG0W0.f90:1-2) and evRGW (src/GW/evRGW.f90:1-2), so QuAcK's RGW
! (src/GW/RGW.f90:85,102) can call them in place of the CPU routines when
! input/hpc_flags asks for the GPU (src/QuAcK/read_hpc_flags.f90:31-33).
! The screening, excitation densities, self-energy and QP solve run in
! libquack_amd.so through the C-ABI; the tables and test dumps are written
! in QuAcK's formats (src/GW/print_RG0W0.f90:33-69, print_evRGW.f90:107-152,
! src/GW/RG0W0.f90:299-305, src/GW/evRGW.f90:264-270).
! phBSE (static kernel, second-order kernel, first-order dynamical
! correction) runs on the GPU after the GW step; the linDM, ppBSE and ACFDT
! branches stay on QuAcK's CPU routines: requesting them here stops with a
! nonzero code.

subroutine qk_unsupported(what)
  implicit none
  character(len=*), intent(in) :: what
  write(*,'(A,A,A)') ' quack_amd: ', what, ' is not available on the GPU path; use the CPU routine'
  error stop 75
end subroutine

subroutine qk_print_table(title, nOrb, nC, nO, nR, eHF, SigC, Z, eGW)
  implicit none
  character(len=*), intent(in) :: title
  integer, intent(in) :: nOrb, nC, nO, nR
  double precision, intent(in) :: eHF(nOrb), SigC(nOrb), Z(nOrb), eGW(nOrb)
  double precision, parameter :: HaToeV = 27.211386245988d0
  integer :: p
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,*) title
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(1X,A1,1X,A3,1X,A1,1X,A15,1X,A1,1X,A15,1X,A1,1X,A15,1X,A1,1X,A15,1X,A1,1X)') &
            '|','#','|','e_HF (eV)','|','Sig_GW (eV)','|','Z','|','e_GW (eV)','|'
  write(*,*)'-------------------------------------------------------------------------------'
  do p = nC+1, nOrb-nR
    if(p == nO+1) then
      write(*,*)'-------------------------------------------------------------------------------'
      write(*,*)'-------------------------------------------------------------------------------'
    end if
    write(*,'(1X,A1,1X,I3,1X,A1,1X,F15.6,1X,A1,1X,F15.6,1X,A1,1X,F15.6,1X,A1,1X,F15.6,1X,A1,1X)') &
    '|',p,'|',eHF(p)*HaToeV,'|',SigC(p)*HaToeV,'|',Z(p),'|',eGW(p)*HaToeV,'|'
  end do
  write(*,*)'-------------------------------------------------------------------------------'
end subroutine

subroutine RG0W0_hip(dotest,doACFDT,exchange_kernel,doXBS,dophBSE,dophBSE2,TDA_W,TDA,dBSE,dTDA,doppBSE,singlet,triplet, &
                     linearize,eta,doSRG,do_linDM_GW,nBas,nOrb,nC,nO,nV,nR,nS,ENuc,ERHF,ERI,dipole_int,eHF,eGW_out)
  use quack_amd_gw
  implicit none
  logical, intent(in) :: dotest, doACFDT, exchange_kernel, doXBS, dophBSE, dophBSE2, doppBSE, TDA_W, TDA, dBSE, dTDA
  logical, intent(in) :: singlet, triplet, linearize, doSRG, do_linDM_GW
  double precision, intent(in) :: eta
  integer, intent(in) :: nBas, nOrb, nC, nO, nV, nR, nS
  double precision, intent(in) :: ENuc, ERHF
  double precision, intent(in) :: ERI(nOrb,nOrb,nOrb,nOrb), dipole_int(nOrb,nOrb,3), eHF(nOrb)
  double precision, intent(out) :: eGW_out(nOrb)
  double precision, parameter :: HaToeV = 27.211386245988d0
  double precision, allocatable :: SigC(:), Z(:), Om(:)
  double precision :: EcRPA, EcGM, eHOMO, eLUMO
  type(c_ptr) :: h
  integer(c_int) :: rc

  if(doppBSE) call qk_unsupported('ppBSE@G0W0')
  if(doACFDT) call qk_unsupported('ACFDT@G0W0')
  if(do_linDM_GW) call qk_unsupported('linDM@G0W0')

  write(*,*)
  write(*,*)'*******************************'
  write(*,*)'* Restricted G0W0 Calculation *'
  write(*,*)'*******************************'
  write(*,*)
  if(TDA_W) then
    write(*,*) 'Tamm-Dancoff approximation for dynamical screening!'
    write(*,*)
  end if
  if(doSRG) then
    write(*,*) '*** SRG regularized G0W0 scheme ***'
    write(*,*)
  end if

  allocate(SigC(nOrb), Z(nOrb), Om(nS))
  rc = qk_gw_create(h, int(nOrb,c_int), int(nOrb,c_int), int(nC,c_int), int(nO,c_int), int(nV,c_int), &
                    int(nR,c_int), int(nS,c_int))
  call qk_check(rc, c_null_ptr, 'qk_gw_create')
  rc = qk_gw_set_eri_mo(h, ERI)
  call qk_check(rc, h, 'qk_gw_set_eri_mo')
  rc = qk_rg0w0(h, eHF, eta, merge(1_c_int, 0_c_int, linearize), merge(1_c_int, 0_c_int, TDA_W), &
                merge(1_c_int, 0_c_int, doSRG), eGW_out, SigC, Z, Om, EcRPA, EcGM)
  call qk_check(rc, h, 'qk_rg0w0')
  rc = qk_gw_destroy(h)

  if(linearize) then
    write(*,*) ' *** Quasiparticle energies obtained by linearization *** '
  else
    write(*,*) ' *** Quasiparticle energies obtained by root search *** '
  end if
  write(*,*)

  eHOMO = maxval(eGW_out(1:nO))
  eLUMO = minval(eGW_out(nO+1:nOrb))
  call qk_print_table(' G0W0@RHF calculation ', nOrb, nC, nO, nR, eHF, SigC, Z, eGW_out)
  write(*,'(2X,A60,F15.6,A3)') 'G0W0@RHF HOMO      energy = ',eHOMO*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'G0W0@RHF LUMO      energy = ',eLUMO*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'G0W0@RHF HOMO-LUMO gap    = ',(eLUMO-eHOMO)*HaToeV,' eV'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@G0W0@RHF total       energy = ',ENuc + ERHF + EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@G0W0@RHF correlation energy = ',EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@G0W0@RHF total       energy = ',ENuc + ERHF + EcGM,' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@G0W0@RHF correlation energy = ',EcGM,' au'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,*)

  ! phBSE at the G0W0 energies (GW/RG0W0.f90:239-255): W from eHF
  if(dophBSE) call qk_phbse_print('G0W0', dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet, eta, &
                                  nOrb, nC, nO, nV, nR, nS, ENuc, ERHF, ERI, eHF, eGW_out)

  if(dotest) then
    call dump_test_value('R','G0W0 correlation energy',EcRPA)
    call dump_test_value('R','G0W0 HOMO energy',eGW_out(nO))
    call dump_test_value('R','G0W0 LUMO energy',eGW_out(nO+1))
  end if

  deallocate(SigC, Z, Om)
end subroutine

! phBSE (qk_rgw_phbse_dyn) after a GW calculation, with QuAcK's output: the
! excitation tables that RGW_phBSE prints (GW/RGW_phBSE.f90:145,180), the
! second-order kernel's notice (:117-121), the first-order dynamical
! correction's table (GW/RGW_phBSE_dynamic_perturbation.f90:62-122) after
! each manifold, and the Tr@BSE energies of GW/RG0W0.f90:248-255 /
! GW/evRGW.f90:219-226 (Eref: ERHF there, EqsGW after qsGW).
subroutine qk_phbse_print(tag, dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet, eta, nOrb, nC, &
                          nO, nV, nR, nS, ENuc, Eref, ERI, eW, eGW)
  use quack_amd_gw
  implicit none
  character(len=*), intent(in) :: tag
  logical, intent(in) :: dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet
  double precision, intent(in) :: eta, ENuc, Eref
  integer, intent(in) :: nOrb, nC, nO, nV, nR, nS
  double precision, intent(in) :: ERI(nOrb,nOrb,nOrb,nOrb), eW(nOrb), eGW(nOrb)
  double precision, allocatable :: OmBSE(:)
  double precision :: EcBSE(2), OmDyn(20), ZDyn(20)
  integer(c_int) :: rc
  allocate(OmBSE(2*nS))
  rc = qk_rgw_phbse_dyn(merge(1_c_int, 0_c_int, dophBSE2), merge(1_c_int, 0_c_int, exchange_kernel),       &
                        merge(1_c_int, 0_c_int, TDA_W), merge(1_c_int, 0_c_int, TDA),                        &
                        merge(1_c_int, 0_c_int, dBSE), merge(1_c_int, 0_c_int, dTDA),                        &
                        merge(1_c_int, 0_c_int, singlet), merge(1_c_int, 0_c_int, triplet), eta,           &
                        int(nOrb,c_int), int(nC,c_int), int(nO,c_int), int(nV,c_int), int(nR,c_int),      &
                        int(nS,c_int), ERI, eW, eGW, EcBSE, OmBSE, OmDyn, ZDyn)
  call qk_check(rc, c_null_ptr, 'qk_rgw_phbse_dyn')
  if(singlet) then
    if(dophBSE2) then
      write(*,*)
      write(*,*) '*** Second-order BSE static kernel activated! ***'
      write(*,*)
    end if
    call qk_print_excitations('phBSE@GW@RHF', 'singlet', nS, OmBSE(1:nS))
    if(dBSE) call qk_print_dynamic(dophBSE2, dTDA, nS, OmBSE(1:nS), OmDyn(1:10), ZDyn(1:10))
  end if
  if(triplet) then
    call qk_print_excitations('phBSE@GW@RHF', 'triplet', nS, OmBSE(nS+1:2*nS))
    if(dBSE) call qk_print_dynamic(dophBSE2, dTDA, nS, OmBSE(nS+1:2*nS), OmDyn(11:20), ZDyn(11:20))
  end if
  write(*,*)
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A50,F20.10,A3)') 'Tr@BSE@'//tag//'@RHF correlation energy (singlet) = ',EcBSE(1),' au'
  write(*,'(2X,A50,F20.10,A3)') 'Tr@BSE@'//tag//'@RHF correlation energy (triplet) = ',EcBSE(2),' au'
  write(*,'(2X,A50,F20.10,A3)') 'Tr@BSE@'//tag//'@RHF correlation energy           = ',sum(EcBSE),' au'
  write(*,'(2X,A50,F20.10,A3)') 'Tr@BSE@'//tag//'@RHF total       energy           = ',ENuc + Eref + sum(EcBSE),' au'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,*)
  deallocate(OmBSE)
end subroutine

! The table of GW/RGW_phBSE_dynamic_perturbation.f90:56-122 for one manifold.
subroutine qk_print_dynamic(dophBSE2, dTDA, nS, OmBSE, OmDyn, ZDyn)
  implicit none
  logical, intent(in) :: dophBSE2, dTDA
  integer, intent(in) :: nS
  double precision, intent(in) :: OmBSE(nS), OmDyn(10), ZDyn(10)
  double precision, parameter :: HaToeV = 27.211386245988d0
  integer :: ia
  if(dTDA) then
    write(*,*)
    write(*,*) '*** dynamical TDA activated ***'
    write(*,*)
  end if
  if(dophBSE2) then
    write(*,*)
    write(*,*) '*** Second-order BSE dynamic kernel activated! ***'
    write(*,*)
  end if
  write(*,*) '---------------------------------------------------------------------------------------------------'
  write(*,*) ' First-order dynamical correction to static Bethe-Salpeter excitation energies                     '
  write(*,*) '---------------------------------------------------------------------------------------------------'
  write(*,'(2X,A5,1X,A20,1X,A20,1X,A20,1X,A20)') '#','Static (eV)','Dynamic (eV)','Correction (eV)','Renorm. (eV)'
  write(*,*) '---------------------------------------------------------------------------------------------------'
  do ia = 1, min(nS, 10)
    write(*,'(2X,I5,5X,F15.6,5X,F15.6,5X,F15.6,5X,F15.6)') &
      ia, OmBSE(ia)*HaToeV, (OmBSE(ia) + OmDyn(ia))*HaToeV, OmDyn(ia)*HaToeV, ZDyn(ia)
  end do
  write(*,*) '---------------------------------------------------------------------------------------------------'
  write(*,*)
end subroutine

! Excitation energies of one spin manifold, in the table of QuAcK's
! LR/print_excitation_energies.f90 (at most the first 50 states).
subroutine qk_print_excitations(method, manifold, nS, Om)
  implicit none
  character(len=*), intent(in) :: method, manifold
  integer, intent(in) :: nS
  double precision, intent(in) :: Om(nS)
  double precision, parameter :: HaToeV = 27.211386245988d0
  integer :: m
  write(*,*)
  write(*,*)'-------------------------------------------------------------'
  write(*,'(1X,A15,A15,A15,A9)') trim(method),' calculation: ',trim(manifold),' manifold'
  write(*,*)'-------------------------------------------------------------'
  write(*,'(1X,A1,1X,A5,1X,A1,1X,A23,1X,A1,1X,A23,1X,A1,1X)') &
            '|','State','|',' Excitation energy (au) ','|',' Excitation energy (eV) ','|'
  write(*,*)'-------------------------------------------------------------'
  do m = 1, min(50, nS)
    write(*,'(1X,A1,1X,I5,1X,A1,1X,F23.6,1X,A1,1X,F23.6,1X,A1,1X)') '|',m,'|',Om(m),'|',Om(m)*HaToeV,'|'
  end do
  write(*,*)'-------------------------------------------------------------'
  write(*,*)
end subroutine

subroutine evRGW_hip(dotest,maxSCF,thresh,max_diis,doACFDT,exchange_kernel,doXBS,dophBSE,dophBSE2,TDA_W,TDA,dBSE,dTDA, &
                     doppBSE,singlet,triplet,linearize,eta,doSRG,nBas,nOrb,nC,nO,nV,nR,nS,ENuc,ERHF,ERI,dipole_int,eHF)
  use quack_amd_gw
  implicit none
  logical, intent(in) :: dotest
  integer, intent(in) :: maxSCF, max_diis
  double precision, intent(in) :: thresh, ENuc, ERHF
  logical, intent(in) :: doACFDT, exchange_kernel, doXBS, dophBSE, dophBSE2, TDA_W, TDA, dBSE, dTDA, doppBSE
  logical, intent(in) :: singlet, triplet, linearize, doSRG
  double precision, intent(in) :: eta
  integer, intent(in) :: nBas, nOrb, nC, nO, nV, nR, nS
  double precision, intent(in) :: eHF(nOrb), ERI(nOrb,nOrb,nOrb,nOrb), dipole_int(nOrb,nOrb,3)
  double precision, parameter :: HaToeV = 27.211386245988d0
  double precision, allocatable :: eGW(:), SigC(:), Z(:), Om(:)
  double precision :: EcRPA, EcGM
  integer(c_int) :: rc, nSCF
  type(c_ptr) :: h

  if(doppBSE) call qk_unsupported('ppBSE@evGW')
  if(doACFDT) call qk_unsupported('ACFDT@evGW')

  write(*,*)
  write(*,*)'*******************************'
  write(*,*)'* Restricted evGW Calculation *'
  write(*,*)'*******************************'
  write(*,*)
  if(TDA_W) then
    write(*,*) 'Tamm-Dancoff approximation for dynamic screening!'
    write(*,*)
  end if
  if(doSRG) then
    write(*,*) '*** SRG regularized evGW scheme ***'
    write(*,*)
  end if

  allocate(eGW(nOrb), SigC(nOrb), Z(nOrb), Om(nS))
  rc = qk_gw_create(h, int(nOrb,c_int), int(nOrb,c_int), int(nC,c_int), int(nO,c_int), int(nV,c_int), &
                    int(nR,c_int), int(nS,c_int))
  call qk_check(rc, c_null_ptr, 'qk_gw_create')
  rc = qk_gw_set_eri_mo(h, ERI)
  call qk_check(rc, h, 'qk_gw_set_eri_mo')
  rc = qk_evrgw(h, eHF, c_null_ptr, int(maxSCF,c_int), thresh, int(max_diis,c_int), eta, &
                merge(1_c_int, 0_c_int, linearize), merge(1_c_int, 0_c_int, TDA_W), merge(1_c_int, 0_c_int, doSRG), &
                eGW, SigC, Z, Om, EcRPA, EcGM, nSCF)
  if(rc == QK_E_NOCONV) then
    write(*,*)
    write(*,*)'!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!'
    write(*,*)'                 Convergence failed                 '
    write(*,*)'!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!'
    write(*,*)
    error stop 74
  end if
  call qk_check(rc, h, 'qk_evrgw')
  rc = qk_gw_destroy(h)

  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(1X,A20,I0,A1,I0,A16)')' Self-consistent evG',nSCF-1,'W',nSCF-1,'@RHF calculation'
  call qk_print_table('', nOrb, nC, nO, nR, eHF, SigC, Z, eGW)
  write(*,'(2X,A10,I3)')   'Iteration ',nSCF-1
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') 'evGW@RHF HOMO      energy = ',eGW(nO)*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'evGW@RHF LUMO      energy = ',eGW(nO+1)*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'evGW@RHF HOMO-LUMO gap    = ',(eGW(nO+1)-eGW(nO))*HaToeV,' eV'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@evGW@RHF total       energy = ',ENuc + ERHF + EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@evGW@RHF correlation energy = ',EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@evGW@RHF total       energy = ',ENuc + ERHF + EcGM,' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@evGW@RHF correlation energy = ',EcGM,' au'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,*)

  ! phBSE at the converged energies (GW/evRGW.f90:214-226): W from eGW too
  if(dophBSE) call qk_phbse_print('evGW', dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet, eta, &
                                  nOrb, nC, nO, nV, nR, nS, ENuc, ERHF, ERI, eGW, eGW)

  if(dotest) then
    call dump_test_value('R','evGW correlation energy',EcRPA)
    call dump_test_value('R','evGW HOMO energy',eGW(nO))
    call dump_test_value('R','evGW LUMO energy',eGW(nO+1))
  end if

  deallocate(eGW, SigC, Z, Om)
end subroutine
