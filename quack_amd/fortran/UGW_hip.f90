! Drop-in GPU forms of QuAcK's unrestricted G0W0 and evGW drivers.
!
! UG0W0_hip and evUGW_hip take exactly the argument lists of UG0W0
! (src/GW/UG0W0.f90:1-3) and evUGW (src/GW/evUGW.f90:1-3), so QuAcK's UGW
! (src/GW/UGW.f90:82-84,114-116) can call them in place of the CPU routines.
! Screening, excitation densities, self-energies and QP solves run in
! libquack_amd.so (qk_ug0w0, qk_evugw); the summaries and test dumps follow
! src/GW/print_UG0W0.f90:55-63, print_evUGW.f90:70-81, UG0W0.f90:227-233 and
! evUGW.f90:279-285.  Frozen core and Rydberg orbitals (nC, nR per spin) and
! the static phBSE after the GW step run on the GPU; the dynamical BSE
! correction and ACFDT stay on QuAcK's CPU routines: requesting them here
! stops with a nonzero code.

subroutine qk_print_utable(title, nBas, nO, eHF, SigC, Z, eGW)
  implicit none
  character(len=*), intent(in) :: title
  integer, intent(in) :: nBas, nO(2)
  double precision, intent(in) :: eHF(nBas,2), SigC(nBas,2), Z(nBas,2), eGW(nBas,2)
  integer :: is
  do is = 1, 2
    if(is == 1) call qk_print_table(title//'(spin-up)', nBas, 0, nO(is), 0, eHF(:,is), SigC(:,is), Z(:,is), eGW(:,is))
    if(is == 2) call qk_print_table(title//'(spin-down)', nBas, 0, nO(is), 0, eHF(:,is), SigC(:,is), Z(:,is), &
                                    eGW(:,is))
  end do
end subroutine

subroutine qk_ugw_checks(doACFDT, dophBSE, dBSE)
  implicit none
  logical, intent(in) :: doACFDT, dophBSE, dBSE
  if(dophBSE .and. dBSE) call qk_unsupported('dynamical phBSE@UGW')
  if(doACFDT) call qk_unsupported('ACFDT@UGW')
end subroutine

! Static phBSE after an unrestricted GW calculation (qk_ugw_phbse,
! GW/UGW_phBSE.f90:60-196) with QuAcK's output: the excitation tables of the
! spin-conserved and spin-flip manifolds (:128,:170) and the Tr@BSE lines of
! GW/UG0W0.f90:195-202 (label width 60) / GW/evUGW.f90:248-255 (width 50).
subroutine qk_uphbse_print(tag, wid, exchange_kernel, TDA_W, TDA, spin_conserved, spin_flip, eta, nBas, nC, nO, nR, &
                           ENuc, EUHF, ERI_aaaa, ERI_aabb, ERI_bbbb, eW, eGW)
  use quack_amd_gw
  implicit none
  character(len=*), intent(in) :: tag
  integer, intent(in) :: wid, nBas, nC(2), nO(2), nR(2)
  logical, intent(in) :: exchange_kernel, TDA_W, TDA, spin_conserved, spin_flip
  double precision, intent(in) :: eta, ENuc, EUHF, eW(nBas,2), eGW(nBas,2)
  double precision, intent(in) :: ERI_aaaa(nBas,nBas,nBas,nBas), ERI_aabb(nBas,nBas,nBas,nBas)
  double precision, intent(in) :: ERI_bbbb(nBas,nBas,nBas,nBas)
  double precision, allocatable :: OmSC(:), OmSF(:)
  double precision :: EcBSE(2)
  integer :: nSC, nSF
  integer(c_int) :: rc
  character(len=80) :: fmt
  nSC = (nO(1) - nC(1))*(nBas - nR(1) - nO(1)) + (nO(2) - nC(2))*(nBas - nR(2) - nO(2))
  nSF = (nO(1) - nC(1))*(nBas - nR(2) - nO(2)) + (nO(2) - nC(2))*(nBas - nR(1) - nO(1))
  allocate(OmSC(nSC), OmSF(nSF))
  if(TDA) then
    write(*,*) 'Tamm-Dancoff approximation activated in phBSE!'
    write(*,*)
  end if
  rc = qk_ugw_phbse(int(nBas,c_int), int(nC,c_int), int(nO,c_int), int(nR,c_int), ERI_aaaa, ERI_aabb, ERI_bbbb, &
                    eW, eGW, eta, merge(1_c_int, 0_c_int, TDA_W), merge(1_c_int, 0_c_int, TDA),                  &
                    merge(1_c_int, 0_c_int, spin_conserved), merge(1_c_int, 0_c_int, spin_flip),                &
                    merge(1_c_int, 0_c_int, exchange_kernel), EcBSE, OmSC, OmSF)
  call qk_check(rc, c_null_ptr, 'qk_ugw_phbse')
  if(spin_conserved) call qk_print_excitations('phBSE@GW@UHF', 'spin-conserved', nSC, OmSC)
  if(spin_flip) call qk_print_excitations('phBSE@GW@UHF', 'spin-flip', nSF, OmSF)
  write(fmt,'(A,I0,A)') '(2X,A', wid, ',F20.10,A3)'
  write(*,*)
  write(*,*)'-------------------------------------------------------------------------------'
  if(wid == 60) then
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF correlation energy (spin-conserved) = ',EcBSE(1),' au'
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF correlation energy (spin-flip)      = ',EcBSE(2),' au'
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF correlation energy                  = ',sum(EcBSE),' au'
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF total       energy                  = ',ENuc + EUHF + sum(EcBSE),' au'
  else
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF correlation energy (spin-conserved) =',EcBSE(1),' au'
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF correlation energy (spin-flip)      =',EcBSE(2),' au'
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF correlation energy                  =',sum(EcBSE),' au'
    write(*,fmt) 'Tr@BSE@'//tag//'@UHF total energy                        =',ENuc + EUHF + sum(EcBSE),' au'
  end if
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,*)
  deallocate(OmSC, OmSF)
end subroutine

subroutine UG0W0_hip(dotest,doACFDT,exchange_kernel,doXBS,dophBSE,TDA_W,TDA,dBSE,dTDA,spin_conserved,spin_flip, &
                     linearize,eta,doSRG,nBas,nC,nO,nV,nR,nS,ENuc,EUHF,S,ERI_aaaa,ERI_aabb,ERI_bbbb,       &
                     dipole_int_aa,dipole_int_bb,cHF,eHF)
  use quack_amd_gw
  implicit none
  logical, intent(in) :: dotest, doACFDT, exchange_kernel, doXBS, dophBSE, TDA_W, TDA, dBSE, dTDA
  logical, intent(in) :: spin_conserved, spin_flip, linearize, doSRG
  double precision, intent(in) :: eta, ENuc, EUHF
  integer, intent(in) :: nBas, nC(2), nO(2), nV(2), nR(2), nS(2)
  double precision, intent(in) :: S(nBas,nBas), cHF(nBas,nBas,2), eHF(nBas,2)
  double precision, intent(in) :: ERI_aaaa(nBas,nBas,nBas,nBas), ERI_aabb(nBas,nBas,nBas,nBas)
  double precision, intent(in) :: ERI_bbbb(nBas,nBas,nBas,nBas)
  double precision, intent(in) :: dipole_int_aa(nBas,nBas,3), dipole_int_bb(nBas,nBas,3)
  double precision, parameter :: HaToeV = 27.211386245988d0
  double precision, allocatable :: eGW(:,:), SigC(:,:), Z(:,:), Om(:)
  double precision :: EcRPA, EcGM(2), eHOMO, eLUMO
  integer(c_int) :: rc

  call qk_ugw_checks(doACFDT, dophBSE, dBSE)
  write(*,*)
  write(*,*)'*********************************'
  write(*,*)'* Unrestricted G0W0 Calculation *'
  write(*,*)'*********************************'
  write(*,*)
  if(TDA_W) write(*,*) 'Tamm-Dancoff approximation for dynamic screening!'
  if(doSRG) write(*,*) '*** SRG regularized G0W0 scheme ***'

  allocate(eGW(nBas,2), SigC(nBas,2), Z(nBas,2), Om(nS(1)+nS(2)))
  rc = qk_ug0w0(int(nBas,c_int), int(nC,c_int), int(nO,c_int), int(nR,c_int), ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, eta, &
                merge(1_c_int, 0_c_int, linearize), merge(1_c_int, 0_c_int, TDA_W), merge(1_c_int, 0_c_int, doSRG), &
                eGW, SigC, Z, Om, EcRPA, EcGM)
  call qk_check(rc, c_null_ptr, 'qk_ug0w0')

  call qk_print_utable(' G0W0@UHF calculation ', nBas, nO, eHF, SigC, Z, eGW)
  eHOMO = max(eGW(nO(1),1), eGW(nO(2),2))
  eLUMO = min(eGW(nO(1)+1,1), eGW(nO(2)+1,2))
  write(*,'(2X,A60,F15.6,A3)') 'G0W0@UHF HOMO      energy = ',eHOMO*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'G0W0@UHF LUMO      energy = ',eLUMO*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'G0W0@UHF HOMO-LUMO gap    = ',(eLUMO-eHOMO)*HaToeV,' eV'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@G0W0@UHF total       energy = ',ENuc + EUHF + EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@G0W0@UHF correlation energy = ',EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@G0W0@UHF total       energy = ',ENuc + EUHF + sum(EcGM),' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@G0W0@UHF correlation energy = ',sum(EcGM),' au'
  write(*,*)'-------------------------------------------------------------------------------'

  ! static phBSE at the G0W0 energies, W from eHF (GW/UG0W0.f90:190-202)
  if(dophBSE) call qk_uphbse_print('G0W0', 60, exchange_kernel, TDA_W, TDA, spin_conserved, spin_flip, eta, nBas, &
                                   nC, nO, nR, ENuc, EUHF, ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, eGW)

  if(dotest) then
    call dump_test_value('U','G0W0 correlation energy',EcRPA)
    call dump_test_value('U','G0W0 HOMOa energy',eGW(nO(1),1))
    call dump_test_value('U','G0W0 LUMOa energy',eGW(nO(1)+1,1))
    call dump_test_value('U','G0W0 HOMOa energy',eGW(nO(2),2))
    call dump_test_value('U','G0W0 LUMOa energy',eGW(nO(2)+1,2))
  end if
  deallocate(eGW, SigC, Z, Om)
end subroutine

subroutine evUGW_hip(dotest,maxSCF,thresh,max_diis,doACFDT,exchange_kernel,doXBS,BSE,TDA_W,TDA,dBSE,dTDA, &
                     spin_conserved,spin_flip,linearize,eta,doSRG,nBas,nC,nO,nV,nR,nS,ENuc,   &
                     EUHF,S,ERI_aaaa,ERI_aabb,ERI_bbbb,dipole_int_aa,dipole_int_bb,cHF,eHF)
  use quack_amd_gw
  implicit none
  logical, intent(in) :: dotest, doACFDT, exchange_kernel, doXBS, BSE, TDA_W, TDA, dBSE, dTDA
  logical, intent(in) :: spin_conserved, spin_flip, linearize, doSRG
  integer, intent(in) :: maxSCF, max_diis
  double precision, intent(in) :: thresh, eta, ENuc, EUHF
  integer, intent(in) :: nBas, nC(2), nO(2), nV(2), nR(2), nS(2)
  double precision, intent(in) :: S(nBas,nBas), cHF(nBas,nBas,2), eHF(nBas,2)
  double precision, intent(in) :: ERI_aaaa(nBas,nBas,nBas,nBas), ERI_aabb(nBas,nBas,nBas,nBas)
  double precision, intent(in) :: ERI_bbbb(nBas,nBas,nBas,nBas)
  double precision, intent(in) :: dipole_int_aa(nBas,nBas,3), dipole_int_bb(nBas,nBas,3)
  double precision, parameter :: HaToeV = 27.211386245988d0
  double precision, allocatable :: eGW(:,:), SigC(:,:), Z(:,:), Om(:)
  double precision :: EcRPA, EcGM(2), eHOMO, eLUMO
  integer(c_int) :: rc, nSCF

  call qk_ugw_checks(doACFDT, BSE, dBSE)
  write(*,*)
  write(*,*)'*********************************'
  write(*,*)'| Unrestricted evGW Calculation *'
  write(*,*)'*********************************'
  write(*,*)
  if(TDA_W) write(*,*) 'Tamm-Dancoff approximation for dynamic screening!'
  if(doSRG) write(*,*) '*** SRG regularized evGW scheme ***'

  allocate(eGW(nBas,2), SigC(nBas,2), Z(nBas,2), Om(nS(1)+nS(2)))
  rc = qk_evugw(int(nBas,c_int), int(nC,c_int), int(nO,c_int), int(nR,c_int), ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, &
                int(maxSCF,c_int), thresh, int(max_diis,c_int), eta, merge(1_c_int, 0_c_int, linearize), &
                merge(1_c_int, 0_c_int, TDA_W), merge(1_c_int, 0_c_int, doSRG), eGW, SigC, Z, Om, EcRPA, EcGM, nSCF)
  if(rc == QK_E_NOCONV) then
    ! GW/evUGW.f90:220-228
    write(*,*)
    write(*,*)'!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!'
    write(*,*)'                 Convergence failed                 '
    write(*,*)'!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!'
    write(*,*)
    stop
  end if
  call qk_check(rc, c_null_ptr, 'qk_evugw')

  call qk_print_utable(' evGW@UHF calculation ', nBas, nO, eHF, SigC, Z, eGW)
  write(*,'(2X,A14,I15)') 'Cycles       = ', nSCF
  eHOMO = max(eGW(nO(1),1), eGW(nO(2),2))
  eLUMO = min(eGW(nO(1)+1,1), eGW(nO(2)+1,2))
  write(*,'(2X,A60,F15.6,A3)') 'evGW@UHF HOMO      energy = ',eHOMO*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'evGW@UHF LUMO      energy = ',eLUMO*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'evGW@UHF HOMO-LUMO gap    = ',(eLUMO-eHOMO)*HaToeV,' eV'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') 'RPA@evG@UHFW total       energy = ',ENuc + EUHF + EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') 'RPA@evG@UHFW correlation energy = ',EcRPA,' au'
  write(*,'(2X,A60,F15.6,A3)') ' GM@evG@UHFW total       energy = ',ENuc + EUHF + sum(EcGM),' au'
  write(*,'(2X,A60,F15.6,A3)') ' GM@evG@UHFW correlation energy = ',sum(EcGM),' au'
  write(*,*)'-------------------------------------------------------------------------------'

  ! static phBSE at the converged energies, W from eGW too (GW/evUGW.f90:243-255)
  if(BSE) call qk_uphbse_print('evGW', 50, exchange_kernel, TDA_W, TDA, spin_conserved, spin_flip, eta, nBas, nC, &
                               nO, nR, ENuc, EUHF, ERI_aaaa, ERI_aabb, ERI_bbbb, eGW, eGW)

  if(dotest) then
    call dump_test_value('U','evGW correlation energy',EcRPA)
    call dump_test_value('U','evGW HOMOa energy',eGW(nO(1),1))
    call dump_test_value('U','evGW LUMOa energy',eGW(nO(1)+1,1))
    call dump_test_value('U','evGW HOMOa energy',eGW(nO(2),2))
    call dump_test_value('U','evGW LUMOa energy',eGW(nO(2)+1,2))
  end if
  deallocate(eGW, SigC, Z, Om)
end subroutine
