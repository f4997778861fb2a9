! Drop-in GPU form of QuAcK's quasiparticle self-consistent GW driver.
!
! qsRGW_hip takes the argument list of qsRGW (src/GW/qsRGW.f90:1-4), so
! QuAcK's RGW (src/GW/RGW.f90) can call it in place of the CPU routine.  The
! loop, its stopping rule and the printing stay here as in qsRGW.f90:193-352;
! each pass of the loop body is one call of libquack_amd.so (qk_qsgw_cycle:
! J and K, the MO integrals, screening, the full self-energy, F, DIIS, the new
! orbitals, all on the device).  X is read as nBas x nOrb, the shape QuAcK
! allocates for it.  On return ERI_MO and dipole_int_MO hold the integrals of
! the last cycle's orbitals, as qsRGW leaves them.  linDM, ppBSE and ACFDT
! stay on QuAcK's CPU routines: requesting them here stops with a nonzero
! code; phBSE (static, second-order, dynamical correction) runs on the GPU.

! GW/print_qsRGW.f90:1-150: one cycle's table and energies, and the summary
! once the commutator is below thresh.
subroutine qk_print_qsrgw(nOrb, nC, nO, nR, nSCF, Conv, thresh, eHF, eGW, SigC, Z, ENuc, ET, EV, EJ, EK, EcGM, &
                          EcRPA, EqsGW, dipole)
  implicit none
  integer, intent(in) :: nOrb, nC, nO, nR, nSCF
  double precision, intent(in) :: Conv, thresh, eHF(nOrb), eGW(nOrb), SigC(nOrb,nOrb), Z(nOrb)
  double precision, intent(in) :: ENuc, ET, EV, EJ, EK, EcGM, EcRPA, EqsGW, dipole(3)
  double precision, parameter :: HaToeV = 27.211386245988d0, auToD = 2.5415802529d0
  integer :: p
  write(*,*)'-------------------------------------------------------------------------------'
  if(nSCF < 10) then
    write(*,'(1X,A20,I1,A1,I1,A16)')' Self-consistent qsG',nSCF,'W',nSCF,'@RHF calculation'
  elseif(nSCF < 100) then
    write(*,'(1X,A20,I2,A1,I2,A16)')' Self-consistent qsG',nSCF,'W',nSCF,'@RHF calculation'
  else
    write(*,'(1X,A20,I3,A1,I3,A16)')' Self-consistent qsG',nSCF,'W',nSCF,'@RHF calculation'
  end if
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
    '|',p,'|',eHF(p)*HaToeV,'|',SigC(p,p)*HaToeV,'|',Z(p),'|',eGW(p)*HaToeV,'|'
  end do
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A10,I3)')   'Iteration ',nSCF
  write(*,'(2X,A14,F15.5)')'Convergence = ',Conv
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') 'qsGW@RHF HOMO      energy = ',eGW(nO)*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'qsGW@RHF LUMO      energy = ',eGW(nO+1)*HaToeV,' eV'
  write(*,'(2X,A60,F15.6,A3)') 'qsGW@RHF HOMO-LUMO gap    = ',(eGW(nO+1)-eGW(nO))*HaToeV,' eV'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,'(2X,A60,F15.6,A3)') '      qsGW@RHF total       energy = ',ENuc + EqsGW,' au'
  write(*,'(2X,A60,F15.6,A3)') '      qsGW@RHF exchange    energy = ',EK,' au'
  write(*,'(2X,A60,F15.6,A3)') '   GM@qsGW@RHF correlation energy = ',EcGM,' au'
  write(*,'(2X,A60,F15.6,A3)') 'phRPA@qsGW@RHF correlation energy = ',EcRPA,' au'
  write(*,*)'-------------------------------------------------------------------------------'
  write(*,*)
  if(Conv < thresh) then
    write(*,*)
    write(*,'(A50)')           '---------------------------------------'
    write(*,'(A33)')           ' Summary               '
    write(*,'(A50)')           '---------------------------------------'
    write(*,'(A33,1X,F16.10,A3)') ' One-electron energy = ',ET + EV,' au'
    write(*,'(A33,1X,F16.10,A3)') ' Kinetic      energy = ',ET,' au'
    write(*,'(A33,1X,F16.10,A3)') ' Potential    energy = ',EV,' au'
    write(*,'(A50)')           '---------------------------------------'
    write(*,'(A33,1X,F16.10,A3)') ' Two-electron energy = ',EJ + EK,' au'
    write(*,'(A33,1X,F16.10,A3)') ' Hartree      energy = ',EJ,' au'
    write(*,'(A33,1X,F16.10,A3)') ' Exchange     energy = ',EK,' au'
    write(*,'(A33,1X,F16.10,A3)') ' Correlation  energy = ',EcGM,' au'
    write(*,'(A50)')           '---------------------------------------'
    write(*,'(A33,1X,F16.10,A3)') ' Electronic   energy = ',EqsGW,' au'
    write(*,'(A33,1X,F16.10,A3)') ' Nuclear   repulsion = ',ENuc,' au'
    write(*,'(A33,1X,F16.10,A3)') ' qsRGW        energy = ',ENuc + EqsGW,' au'
    write(*,'(A50)')           '---------------------------------------'
    write(*,'(A36)')           ' Dipole moment (Debye)    '
    write(*,'(10X,4A10)')      'X','Y','Z','Tot.'
    write(*,'(10X,4F10.4)')    (dipole(p)*auToD,p=1,3),norm2(dipole)*auToD
    write(*,'(A50)')           '---------------------------------------'
    write(*,*)
    write(*,'(A50)') '---------------------------------------'
    write(*,'(A50)') ' Restricted qsGW orbital energies (au) '
    write(*,'(A50)') '---------------------------------------'
    do p = 1, nOrb
      write(*,'(1X,I5,1X,F16.10)') p, eGW(p)
    end do
    write(*,*)
  end if
end subroutine

subroutine qsRGW_hip(dotest,maxSCF,thresh,max_diis,doACFDT,exchange_kernel,doXBS,dophBSE,dophBSE2, &
                     TDA_W,TDA,dBSE,dTDA,doppBSE,singlet,triplet,eta,doSRG,do_linDM_GW,            &
                     nNuc,ZNuc,rNuc,ENuc,nBas,nOrb,nC,nO,nV,nR,nS,ERHF,S,X,T,V,Hc,ERI_AO,          &
                     ERI_MO,dipole_int_AO,dipole_int_MO,PHF,cHF,eHF)
  use quack_amd_gw
  implicit none
  logical, intent(in) :: dotest, doACFDT, exchange_kernel, doXBS, dophBSE, dophBSE2, TDA_W, TDA, dBSE, dTDA
  logical, intent(in) :: doppBSE, singlet, triplet, doSRG, do_linDM_GW
  integer, intent(in) :: maxSCF, max_diis, nNuc, nBas, nOrb, nC, nO, nV, nR, nS
  double precision, intent(in) :: thresh, eta, ENuc, ERHF
  double precision, intent(in) :: ZNuc(nNuc), rNuc(nNuc,3)
  double precision, intent(in) :: S(nBas,nBas), X(nBas,nOrb), T(nBas,nBas), V(nBas,nBas), Hc(nBas,nBas)
  double precision, intent(in) :: ERI_AO(nBas,nBas,nBas,nBas), dipole_int_AO(nBas,nBas,3)
  double precision, intent(inout) :: ERI_MO(nOrb,nOrb,nOrb,nOrb), dipole_int_MO(nOrb,nOrb,3)
  double precision, intent(in) :: PHF(nBas,nBas), cHF(nBas,nOrb), eHF(nOrb)
  double precision, allocatable :: eGW(:), c(:,:), cOld(:,:), SigC(:,:), Z(:), P(:,:)
  double precision :: E(5), EcRPA, EcGM, err_max, Conv, dipole(3)
  integer :: nSCF, ixyz, iNuc, mu, nu
  integer(c_int) :: rc
  type(c_ptr) :: h

  if(do_linDM_GW) call qk_unsupported('linDM@qsGW')
  if(doppBSE) call qk_unsupported('ppBSE@qsGW')
  if(doACFDT) call qk_unsupported('ACFDT@qsGW')

  write(*,*)
  write(*,*)'*******************************'
  write(*,*)'* Restricted qsGW Calculation *'
  write(*,*)'*******************************'
  write(*,*)
  write(*,*) '!! ERIs in MO basis will be overwritten in qsGW !!'
  write(*,*)
  if(TDA_W) then
    write(*,*) 'Tamm-Dancoff approximation for dynamical screening!'
    write(*,*)
  end if
  if(doSRG) then
    write(*,*) '*** SRG regularized qsGW scheme ***'
    write(*,*)
  end if

  allocate(eGW(nOrb), c(nBas,nOrb), cOld(nBas,nOrb), SigC(nOrb,nOrb), Z(nOrb), P(nBas,nBas))
  rc = qk_qsgw_create(h, int(nBas,c_int), int(nOrb,c_int), int(nC,c_int), int(nO,c_int), int(nV,c_int), &
                      int(nR,c_int), int(nS,c_int), int(max_diis,c_int), S, T, V, Hc, X, ERI_AO, PHF, cHF, eHF)
  call qk_check(rc, c_null_ptr, 'qk_qsgw_create')

  ! GW/qsRGW.f90:178-193: the cycle counter starts at -1, Conv at 1
  nSCF = -1
  Conv = 1d0
  eGW(:) = eHF(:)
  c(:,:) = cHF(:,:)
  do while(Conv > thresh .and. nSCF < maxSCF)
    nSCF = nSCF + 1
    cOld(:,:) = c(:,:)
    rc = qk_qsgw_cycle(h, eta, merge(1_c_int, 0_c_int, TDA_W), merge(1_c_int, 0_c_int, doSRG), eGW, c, SigC, Z, &
                       E, EcRPA, EcGM, err_max)
    call qk_check(rc, c_null_ptr, 'qk_qsgw_cycle')
    if(nSCF > 1) Conv = err_max
    ! density of the new orbitals (:339) and the dipole moment (HF/dipole_moment.f90)
    P(:,:) = 2d0*matmul(c(:,1:nO), transpose(c(:,1:nO)))
    do ixyz = 1, 3
      dipole(ixyz) = 0d0
      do iNuc = 1, nNuc
        dipole(ixyz) = dipole(ixyz) + ZNuc(iNuc)*rNuc(iNuc,ixyz)
      end do
      do nu = 1, nBas
        do mu = 1, nBas
          dipole(ixyz) = dipole(ixyz) - P(mu,nu)*dipole_int_AO(mu,nu,ixyz)
        end do
      end do
    end do
    call qk_print_qsrgw(nOrb, nC, nO, nR, nSCF, Conv, thresh, eHF, eGW, SigC, Z, ENuc, E(1), E(2), E(3), E(4), &
                        EcGM, EcRPA, E(5), dipole)
  end do
  rc = qk_qsgw_destroy(h)

  ! GW/qsRGW.f90:354-362
  if(nSCF == maxSCF) then
    write(*,*)
    write(*,*)'!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!'
    write(*,*)'                 Convergence failed                 '
    write(*,*)'!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!!'
    write(*,*)
    error stop 74
  end if

  ! the MO integrals and dipoles of the last cycle's orbitals (:206-210)
  rc = qk_aotomo_eri_rhf(int(nBas,c_int), int(nOrb,c_int), cOld, ERI_AO, ERI_MO)
  call qk_check(rc, c_null_ptr, 'qk_aotomo_eri_rhf')
  do ixyz = 1, 3
    dipole_int_MO(:,:,ixyz) = matmul(transpose(cOld), matmul(dipole_int_AO(:,:,ixyz), cOld))
  end do

  ! phBSE@qsGW (:373-384), with the total energy on EqsGW
  if(dophBSE) call qk_phbse_print('qsGW', dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet, eta, &
                                  nOrb, nC, nO, nV, nR, nS, ENuc, E(5), ERI_MO, eGW, eGW)

  if(dotest) then
    call dump_test_value('R','qsGW correlation energy',EcRPA)
    call dump_test_value('R','qsGW HOMO energy',eGW(nO))
    call dump_test_value('R','qsGW LUMO energy',eGW(nO+1))
  end if
  deallocate(eGW, c, cOld, SigC, Z, P)
end subroutine
