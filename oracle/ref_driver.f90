! TEST INFRASTRUCTURE ONLY -- never part of the product path.
!
! bind(C) entry points over QuAcK's own restricted G0W0/evGW routines, which
! oracle/Makefile compiles unchanged from the QuAcK sources (flang + MKL).
! The product library never links this; tests, __graft_entry__.smoke() and the
! CPU-baseline leg of bench.py load it through ctypes as the checker.
!
! qkref_rg0w0 replays the screening / self-energy / QP sequence of
! GW/RG0W0.f90:88-233 (no linDM, BSE, ACFDT or ppBSE branch) and
! qkref_evrgw the loop of GW/evRGW.f90:104-200, returning a status instead of
! `stop` on non-convergence.  All arrays are Fortran column-major, exactly as
! QuAcK holds them.

subroutine qkref_aotomo(nBas, nOrb, c, ERI_AO, ERI_MO) bind(C, name='qkref_aotomo')
  use iso_c_binding
  implicit none
  integer(c_int), value :: nBas, nOrb
  real(c_double), intent(in) :: c(nBas, nOrb), ERI_AO(nBas, nBas, nBas, nBas)
  real(c_double), intent(out) :: ERI_MO(nOrb, nOrb, nOrb, nOrb)
  call AOtoMO_ERI_RHF(nBas, nOrb, c, ERI_AO, ERI_MO)
end subroutine

subroutine qkref_phrlr(tda, nOrb, nC, nO, nV, nR, nS, e, ERI, EcRPA, Om, XpY, XmY) bind(C, name='qkref_phrlr')
  use iso_c_binding
  implicit none
  integer(c_int), value :: tda, nOrb, nC, nO, nV, nR, nS
  real(c_double), intent(in) :: e(nOrb), ERI(nOrb, nOrb, nOrb, nOrb)
  real(c_double), intent(out) :: EcRPA, Om(nS), XpY(nS, nS), XmY(nS, nS)
  double precision, allocatable :: Aph(:,:), Bph(:,:)
  logical :: TDA_W
  TDA_W = tda /= 0
  allocate(Aph(nS, nS), Bph(nS, nS))
  Bph = 0d0
  call phRLR_A(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, e, ERI, Aph)
  if(.not. TDA_W) call phRLR_B(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
  call phRLR(TDA_W, nS, Aph, Bph, EcRPA, Om, XpY, XmY)
  flush(6)
  deallocate(Aph, Bph)
end subroutine

subroutine qkref_excitation_density(nOrb, nC, nO, nR, nS, ERI, XpY, rho) bind(C, name='qkref_excitation_density')
  use iso_c_binding
  implicit none
  integer(c_int), value :: nOrb, nC, nO, nR, nS
  real(c_double), intent(in) :: ERI(nOrb, nOrb, nOrb, nOrb), XpY(nS, nS)
  real(c_double), intent(out) :: rho(nOrb, nOrb, nS)
  call RGW_excitation_density(nOrb, nC, nO, nR, nS, ERI, XpY, rho)
end subroutine

subroutine qkref_self_energy_diag(srg, eta, nOrb, nC, nO, nV, nR, nS, e, Om, rho, EcGM, SigC, Z) &
    bind(C, name='qkref_self_energy_diag')
  use iso_c_binding
  implicit none
  integer(c_int), value :: srg, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: eta
  real(c_double), intent(in) :: e(nOrb), Om(nS), rho(nOrb, nOrb, nS)
  real(c_double), intent(out) :: EcGM, SigC(nOrb), Z(nOrb)
  if(srg /= 0) then
    call RGW_SRG_self_energy_diag(500d0, nOrb, nOrb, nC, nO, nV, nR, nS, e, Om, rho, EcGM, SigC, Z)
  else
    call RGW_self_energy_diag(eta, nOrb, nOrb, nC, nO, nV, nR, nS, e, Om, rho, EcGM, SigC, Z)
  end if
end subroutine

subroutine qkref_qp_graph(srg, eta, nOrb, nC, nO, nV, nR, nS, eHF, Om, rho, eGWlin, eOld, eGW, Z) &
    bind(C, name='qkref_qp_graph')
  use iso_c_binding
  implicit none
  integer(c_int), value :: srg, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: eta
  real(c_double), intent(in) :: eHF(nOrb), Om(nS), rho(nOrb, nOrb, nS), eGWlin(nOrb), eOld(nOrb)
  real(c_double), intent(out) :: eGW(nOrb)
  real(c_double), intent(inout) :: Z(nOrb)
  call RGW_QP_graph(srg /= 0, eta, 500d0, nOrb, nC, nO, nV, nR, nS, eHF, Om, rho, eGWlin, eOld, eGW, Z)
  flush(6)
end subroutine

! G0W0: GW/RG0W0.f90:88-233.  Om0 is the screening at eHF; EcRPA comes from
! the second RPA at eGW (RG0W0.f90:222-225).
subroutine qkref_rg0w0(tda, lin, srg, eta, nOrb, nC, nO, nV, nR, nS, ENuc, ERHF, ERI, eHF, &
                       eGW, SigC, Z, Om0, EcRPA, EcGM) bind(C, name='qkref_rg0w0')
  use iso_c_binding
  implicit none
  integer(c_int), value :: tda, lin, srg, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: eta, ENuc, ERHF
  real(c_double), intent(in) :: ERI(nOrb, nOrb, nOrb, nOrb), eHF(nOrb)
  real(c_double), intent(out) :: eGW(nOrb), SigC(nOrb), Z(nOrb), Om0(nS), EcRPA, EcGM
  double precision, allocatable :: Aph(:,:), Bph(:,:), Om(:), XpY(:,:), XmY(:,:), rho(:,:,:), eGWlin(:)
  logical :: TDA_W, doSRG

  TDA_W = tda /= 0
  doSRG = srg /= 0
  allocate(Aph(nS,nS), Bph(nS,nS), Om(nS), XpY(nS,nS), XmY(nS,nS), rho(nOrb,nOrb,nS), eGWlin(nOrb))
  Bph = 0d0

                 call phRLR_A(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, eHF, ERI, Aph)
  if(.not.TDA_W) call phRLR_B(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
  call phRLR(TDA_W, nS, Aph, Bph, EcRPA, Om, XpY, XmY)
  Om0(:) = Om(:)

  call RGW_excitation_density(nOrb, nC, nO, nR, nS, ERI, XpY, rho)

  if(doSRG) then
    call RGW_SRG_self_energy_diag(500d0, nOrb, nOrb, nC, nO, nV, nR, nS, eHF, Om, rho, EcGM, SigC, Z)
  else
    call RGW_self_energy_diag(eta, nOrb, nOrb, nC, nO, nV, nR, nS, eHF, Om, rho, EcGM, SigC, Z)
  end if

  eGWlin(:) = eHF(:) + Z(:)*SigC(:)
  if(lin /= 0) then
    eGW(:) = eGWlin(:)
  else
    call RGW_QP_graph(doSRG, eta, 500d0, nOrb, nC, nO, nV, nR, nS, eHF, Om, rho, eGWlin, eHF, eGW, Z)
  end if

                 call phRLR_A(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, eGW, ERI, Aph)
  if(.not.TDA_W) call phRLR_B(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
  call phRLR(TDA_W, nS, Aph, Bph, EcRPA, Om, XpY, XmY)

  call print_RG0W0(nOrb, nC, nO, nV, nR, eHF, ENuc, ERHF, SigC, Z, eGW, EcRPA, EcGM)
  flush(6)
  deallocate(Aph, Bph, Om, XpY, XmY, rho, eGWlin)
end subroutine

! evGW: GW/evRGW.f90:104-200.  Returns 0 on convergence, 1 when nSCF reaches
! maxSCF+1 (where QuAcK prints "Convergence failed" and stops).
function qkref_evrgw(tda, lin, srg, maxSCF, thresh, max_diis, eta, nOrb, nC, nO, nV, nR, nS, ENuc, ERHF, ERI, eHF, &
                     eGW, SigC, Z, EcRPA, EcGM, nSCF_out) result(status) bind(C, name='qkref_evrgw')
  use iso_c_binding
  implicit none
  integer(c_int), value :: tda, lin, srg, maxSCF, max_diis, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: thresh, eta, ENuc, ERHF
  real(c_double), intent(in) :: ERI(nOrb, nOrb, nOrb, nOrb), eHF(nOrb)
  real(c_double), intent(out) :: eGW(nOrb), SigC(nOrb), Z(nOrb), EcRPA, EcGM
  integer(c_int), intent(out) :: nSCF_out
  integer(c_int) :: status
  double precision, allocatable :: Aph(:,:), Bph(:,:), Om(:), XpY(:,:), XmY(:,:), rho(:,:,:), eOld(:)
  double precision, allocatable :: error_diis(:,:), e_diis(:,:)
  logical :: TDA_W, doSRG
  integer :: nSCF, n_diis
  double precision :: Conv, rcond

  TDA_W = tda /= 0
  doSRG = srg /= 0
  allocate(Aph(nS,nS), Bph(nS,nS), Om(nS), XpY(nS,nS), XmY(nS,nS), rho(nOrb,nOrb,nS), eOld(nOrb), &
           error_diis(nOrb,max(max_diis,1)), e_diis(nOrb,max(max_diis,1)))
  Bph = 0d0
  nSCF = 0
  n_diis = 0
  Conv = 1d0
  e_diis(:,:) = 0d0
  error_diis(:,:) = 0d0
  eGW(:) = eHF(:)
  eOld(:) = eGW(:)
  Z(:) = 1d0
  rcond = 0d0

  do while(Conv > thresh .and. nSCF <= maxSCF)
                   call phRLR_A(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, eGW, ERI, Aph)
    if(.not.TDA_W) call phRLR_B(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
    call phRLR(TDA_W, nS, Aph, Bph, EcRPA, Om, XpY, XmY)
    call RGW_excitation_density(nOrb, nC, nO, nR, nS, ERI, XpY, rho)
    if(doSRG) then
      call RGW_SRG_self_energy_diag(500d0, nOrb, nOrb, nC, nO, nV, nR, nS, eGW, Om, rho, EcGM, SigC, Z)
    else
      call RGW_self_energy_diag(eta, nOrb, nOrb, nC, nO, nV, nR, nS, eGW, Om, rho, EcGM, SigC, Z)
    end if
    if(lin /= 0) then
      eGW(:) = eHF(:) + SigC(:)
    else
      call RGW_QP_graph(doSRG, eta, 500d0, nOrb, nC, nO, nV, nR, nS, eHF, Om, rho, eOld, eOld, eGW, Z)
    end if
    Conv = maxval(abs(eGW - eOld))
    call print_evRGW(nOrb, nC, nO, nV, nR, nSCF, Conv, eHF, ENuc, ERHF, SigC, Z, eGW, EcRPA, EcGM)
    if(max_diis > 1) then
      n_diis = min(n_diis+1, max_diis)
      call DIIS_extrapolation(rcond, nOrb, nOrb, n_diis, error_diis, e_diis, eGW-eOld, eGW)
    end if
    eOld(:) = eGW(:)
    nSCF = nSCF + 1
  end do

  flush(6)
  nSCF_out = nSCF
  status = 0
  if(nSCF == maxSCF+1) status = 1
  deallocate(Aph, Bph, Om, XpY, XmY, rho, eOld, error_diis, e_diis)
end function

! GW/RGW_phBSE.f90:113-190, static-kernel path (no phBSE2, no dynamical
! correction, no transition vectors), composed of QuAcK's own routines in the
! same order: dRPA screening at eW, excitation densities, static kernels, then
! the singlet (ispin 1) and triplet (ispin 2) BSE at eGW with dRPA = false.
! GW/RGW_phBSE.f90:100-195 with the second-order kernel (dophBSE2) and the
! first-order dynamical correction (dBSE, dTDA), composed of QuAcK's own
! routines in RGW_phBSE's order: the BSE2 kernels are called on KA/KB, which
! hold the static GW kernels (as RGW_phBSE passes KA_sta/KB_sta), and the
! correction repeats the arithmetic of GW/RGW_phBSE_dynamic_perturbation.f90:73-122
! (which only prints) on QuAcK's dynamic kernels.  OmDyn, ZDyn: the first
! min(nS, 10) corrections of the singlet (1:10) and triplet (11:20) manifolds.
subroutine qkref_phbse_full(tda_w, tda, dosing, dotrip, bse2, dbse, dtda, eta, nOrb, nC, nO, nV, nR, nS, ERI, eW, &
                            eGW, EcBSE, OmS, OmT, OmDyn, ZDyn) bind(C, name='qkref_phbse_full')
  use iso_c_binding
  implicit none
  integer(c_int), value :: tda_w, tda, dosing, dotrip, bse2, dbse, dtda, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: eta
  real(c_double), intent(in) :: ERI(nOrb, nOrb, nOrb, nOrb), eW(nOrb), eGW(nOrb)
  real(c_double), intent(out) :: EcBSE(2), OmS(nS), OmT(nS), OmDyn(20), ZDyn(20)
  double precision, allocatable :: Aph(:,:), Bph(:,:), KA(:,:), KB(:,:), OmRPA(:), XpY(:,:), XmY(:,:), rho(:,:,:)
  double precision, allocatable :: W(:,:,:,:)
  double precision :: EcRPA
  logical :: TDAW, TDAB
  integer :: ispin
  TDAW = tda_w /= 0
  TDAB = tda /= 0
  allocate(Aph(nS, nS), Bph(nS, nS), KA(nS, nS), KB(nS, nS), OmRPA(nS), XpY(nS, nS), XmY(nS, nS), &
           rho(nOrb, nOrb, nS), W(nOrb, nOrb, nOrb, nOrb))
  Aph = 0d0; Bph = 0d0; KA = 0d0; KB = 0d0
  EcBSE = 0d0; OmS = 0d0; OmT = 0d0; OmDyn = 0d0; ZDyn = 0d0
  call phRLR_A(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, eW, ERI, Aph)
  if(.not. TDAW) call phRLR_B(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
  call phRLR(TDAW, nS, Aph, Bph, EcRPA, OmRPA, XpY, XmY)
  call RGW_excitation_density(nOrb, nC, nO, nR, nS, ERI, XpY, rho)
  call RGW_phBSE_static_kernel_A(eta, nOrb, nC, nO, nV, nR, nS, 1d0, ERI, OmRPA, rho, KA)
  if(.not. TDAB) call RGW_phBSE_static_kernel_B(eta, nOrb, nC, nO, nV, nR, nS, 1d0, ERI, OmRPA, rho, KB)
  if(bse2 /= 0) call RGW_phBSE_static_kernel(eta, nOrb, nC, nO, nV, nR, nS, 1d0, ERI, OmRPA, rho, W)
  do ispin = 1, 2
    if(ispin == 1 .and. dosing == 0) cycle
    if(ispin == 2 .and. dotrip == 0) cycle
    Bph = 0d0
    call phRLR_A(ispin, .false., nOrb, nC, nO, nV, nR, nS, 1d0, eGW, ERI, Aph)
    if(.not. TDAB) call phRLR_B(ispin, .false., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
    if(ispin == 1 .and. bse2 /= 0) then
      call RGW_phBSE2_static_kernel_A(eta, nOrb, nC, nO, nV, nR, nS, 1d0, eW, W, KA)
      if(.not. TDAB) call RGW_phBSE2_static_kernel_B(eta, nOrb, nC, nO, nV, nR, nS, 1d0, eW, W, KB)
    end if
    Aph = Aph + KA
    if(.not. TDAB) Bph = Bph + KB
    if(ispin == 1) then
      call phRLR(TDAB, nS, Aph, Bph, EcBSE(1), OmS, XpY, XmY)
      if(dbse /= 0) call qkref_dyn(bse2 /= 0, dtda /= 0, eta, nOrb, nC, nO, nV, nR, nS, eGW, OmRPA, rho, W, OmS, &
                                   XpY, XmY, KA, KB, OmDyn(1:10), ZDyn(1:10))
    else
      call phRLR(TDAB, nS, Aph, Bph, EcBSE(2), OmT, XpY, XmY)
      if(dbse /= 0) call qkref_dyn(bse2 /= 0, dtda /= 0, eta, nOrb, nC, nO, nV, nR, nS, eGW, OmRPA, rho, W, OmT, &
                                   XpY, XmY, KA, KB, OmDyn(11:20), ZDyn(11:20))
    end if
  end do
  EcBSE(2) = 3d0 * EcBSE(2)
  flush(6)
  deallocate(Aph, Bph, KA, KB, OmRPA, XpY, XmY, rho, W)
end subroutine

subroutine qkref_dyn(bse2, dtda, eta, nOrb, nC, nO, nV, nR, nS, eGW, OmRPA, rho, W, OmBSE, XpY, XmY, KA_sta, &
                     KB_sta, OmDyn, ZDyn)
  implicit none
  logical, intent(in) :: bse2, dtda
  integer, intent(in) :: nOrb, nC, nO, nV, nR, nS
  double precision, intent(in) :: eta, eGW(nOrb), OmRPA(nS), rho(nOrb,nOrb,nS), W(nOrb,nOrb,nOrb,nOrb)
  double precision, intent(in) :: OmBSE(nS), XpY(nS,nS), XmY(nS,nS), KA_sta(nS,nS), KB_sta(nS,nS)
  double precision, intent(out) :: OmDyn(10), ZDyn(10)
  double precision, allocatable :: X(:), Y(:), KAp(:,:), ZAp(:,:), KAm(:,:), ZAm(:,:), KBd(:,:)
  integer :: ia
  allocate(X(nS), Y(nS), KAp(nS,nS), ZAp(nS,nS), KAm(nS,nS), ZAm(nS,nS), KBd(nS,nS))
  OmDyn = 0d0; ZDyn = 0d0
  do ia = 1, min(nS, 10)
    X(:) = 0.5d0*(XpY(ia,:) + XmY(ia,:))
    Y(:) = 0.5d0*(XpY(ia,:) - XmY(ia,:))
    call RGW_phBSE_dynamic_kernel_A(eta, nOrb, nC, nO, nV, nR, nS, 1d0, eGW, OmRPA, rho, +OmBSE(ia), KAp, ZAp)
    if(bse2) call RGW_phBSE2_dynamic_kernel_A(eta, nOrb, nC, nO, nV, nR, nS, eGW, W, OmBSE(ia), KAp, ZAp)
    if(dtda) then
      ZDyn(ia)  = dot_product(X, matmul(ZAp, X))
      OmDyn(ia) = dot_product(X, matmul(KAp - KA_sta, X))
    else
      call RGW_phBSE_dynamic_kernel_A(eta, nOrb, nC, nO, nV, nR, nS, 1d0, eGW, OmRPA, rho, -OmBSE(ia), KAm, ZAm)
      call RGW_phBSE_dynamic_kernel_B(eta, nOrb, nC, nO, nV, nR, nS, 1d0, eGW, OmRPA, rho, KBd)
      if(bse2) call RGW_phBSE2_dynamic_kernel_B(eta, nOrb, nC, nO, nV, nR, nS, eGW, W, KBd)
      ZDyn(ia)  = dot_product(X, matmul(ZAp, X)) + dot_product(Y, matmul(ZAm, Y))
      OmDyn(ia) = dot_product(X, matmul(KAp - KA_sta, X)) - dot_product(Y, matmul(KAm - KA_sta, Y)) &
                + dot_product(X, matmul(KBd - KB_sta, Y)) - dot_product(Y, matmul(KBd - KB_sta, X))
    end if
    ZDyn(ia)  = 1d0/(1d0 - ZDyn(ia))
    OmDyn(ia) = ZDyn(ia)*OmDyn(ia)
  end do
  deallocate(X, Y, KAp, ZAp, KAm, ZAm, KBd)
end subroutine

subroutine qkref_phbse(tda_w, tda, dosing, dotrip, eta, nOrb, nC, nO, nV, nR, nS, ERI, eW, eGW, EcBSE, OmS, OmT) &
    bind(C, name='qkref_phbse')
  use iso_c_binding
  implicit none
  integer(c_int), value :: tda_w, tda, dosing, dotrip, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: eta
  real(c_double), intent(in) :: ERI(nOrb, nOrb, nOrb, nOrb), eW(nOrb), eGW(nOrb)
  real(c_double), intent(out) :: EcBSE(2), OmS(nS), OmT(nS)
  double precision, allocatable :: Aph(:,:), Bph(:,:), KA(:,:), KB(:,:), OmRPA(:), XpY(:,:), XmY(:,:), rho(:,:,:)
  double precision :: EcRPA
  logical :: TDAW, TDAB
  TDAW = tda_w /= 0
  TDAB = tda /= 0
  allocate(Aph(nS, nS), Bph(nS, nS), KA(nS, nS), KB(nS, nS), OmRPA(nS), XpY(nS, nS), XmY(nS, nS), &
           rho(nOrb, nOrb, nS))
  Aph = 0d0; Bph = 0d0; KA = 0d0; KB = 0d0
  EcBSE = 0d0; OmS = 0d0; OmT = 0d0
  call phRLR_A(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, eW, ERI, Aph)
  if(.not. TDAW) call phRLR_B(1, .true., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
  call phRLR(TDAW, nS, Aph, Bph, EcRPA, OmRPA, XpY, XmY)
  call RGW_excitation_density(nOrb, nC, nO, nR, nS, ERI, XpY, rho)
  call RGW_phBSE_static_kernel_A(eta, nOrb, nC, nO, nV, nR, nS, 1d0, ERI, OmRPA, rho, KA)
  if(.not. TDAB) call RGW_phBSE_static_kernel_B(eta, nOrb, nC, nO, nV, nR, nS, 1d0, ERI, OmRPA, rho, KB)
  if(dosing /= 0) then
    Bph = 0d0
    call phRLR_A(1, .false., nOrb, nC, nO, nV, nR, nS, 1d0, eGW, ERI, Aph)
    if(.not. TDAB) call phRLR_B(1, .false., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
    Aph = Aph + KA
    if(.not. TDAB) Bph = Bph + KB
    call phRLR(TDAB, nS, Aph, Bph, EcBSE(1), OmS, XpY, XmY)
  end if
  if(dotrip /= 0) then
    Bph = 0d0
    call phRLR_A(2, .false., nOrb, nC, nO, nV, nR, nS, 1d0, eGW, ERI, Aph)
    if(.not. TDAB) call phRLR_B(2, .false., nOrb, nC, nO, nV, nR, nS, 1d0, ERI, Bph)
    Aph = Aph + KA
    if(.not. TDAB) Bph = Bph + KB
    call phRLR(TDAB, nS, Aph, Bph, EcBSE(2), OmT, XpY, XmY)
  end if
  EcBSE(2) = 3d0 * EcBSE(2)
  flush(6)
  deallocate(Aph, Bph, KA, KB, OmRPA, XpY, XmY, rho)
end subroutine

! GW/RGW_self_energy.f90 / GW/RGW_SRG_self_energy.f90: the full self-energy
subroutine qkref_self_energy_full(srg, eta, nOrb, nC, nO, nV, nR, nS, e, Om, rho, EcGM, SigC, Z) &
    bind(C, name='qkref_self_energy_full')
  use iso_c_binding
  implicit none
  integer(c_int), value :: srg, nOrb, nC, nO, nV, nR, nS
  real(c_double), value :: eta
  real(c_double), intent(in) :: e(nOrb), Om(nS), rho(nOrb, nOrb, nS)
  real(c_double), intent(out) :: EcGM, SigC(nOrb, nOrb), Z(nOrb)
  if(srg /= 0) then
    call RGW_SRG_self_energy(500d0, nOrb, nOrb, nC, nO, nV, nR, nS, e, Om, rho, EcGM, SigC, Z)
  else
    call RGW_self_energy(eta, nOrb, nOrb, nC, nO, nV, nR, nS, e, Om, rho, EcGM, SigC, Z)
  end if
  flush(6)
end subroutine

! One cycle of GW/qsRGW.f90:193-333 (nBas == nOrb, DIIS history of one, which
! leaves F unchanged), composed of QuAcK's own routines: returns F and the new
! quasiparticle energies.
subroutine qkref_qsgw_cycle(srg, tda_w, eta, nBas, nC, nO, nV, nR, nS, S, T, V, X, c, P, eGW, ERI_AO, F, eNew) &
    bind(C, name='qkref_qsgw_cycle')
  use iso_c_binding
  implicit none
  integer(c_int), value :: srg, tda_w, nBas, nC, nO, nV, nR, nS
  real(c_double), value :: eta
  real(c_double), intent(in) :: S(nBas,nBas), T(nBas,nBas), V(nBas,nBas), X(nBas,nBas), c(nBas,nBas), P(nBas,nBas)
  real(c_double), intent(in) :: eGW(nBas), ERI_AO(nBas,nBas,nBas,nBas)
  real(c_double), intent(out) :: F(nBas,nBas), eNew(nBas)
  double precision, allocatable :: J(:,:), K(:,:), ERI_MO(:,:,:,:), Aph(:,:), Bph(:,:), Om(:), XpY(:,:), XmY(:,:)
  double precision, allocatable :: rho(:,:,:), SigC(:,:), SigCp(:,:), Z(:), cp(:,:)
  double precision :: EcRPA, EcGM
  logical :: TDAW
  TDAW = tda_w /= 0
  allocate(J(nBas,nBas), K(nBas,nBas), ERI_MO(nBas,nBas,nBas,nBas), Aph(nS,nS), Bph(nS,nS), Om(nS), &
           XpY(nS,nS), XmY(nS,nS), rho(nBas,nBas,nS), SigC(nBas,nBas), SigCp(nBas,nBas), Z(nBas), cp(nBas,nBas))
  Bph = 0d0
  call Hartree_matrix_AO_basis(nBas, P, ERI_AO, J)
  call exchange_matrix_AO_basis(nBas, P, ERI_AO, K)
  call AOtoMO_ERI_RHF(nBas, nBas, c, ERI_AO, ERI_MO)
  call phRLR_A(1, .true., nBas, nC, nO, nV, nR, nS, 1d0, eGW, ERI_MO, Aph)
  if(.not. TDAW) call phRLR_B(1, .true., nBas, nC, nO, nV, nR, nS, 1d0, ERI_MO, Bph)
  call phRLR(TDAW, nS, Aph, Bph, EcRPA, Om, XpY, XmY)
  call RGW_excitation_density(nBas, nC, nO, nR, nS, ERI_MO, XpY, rho)
  if(srg /= 0) then
    call RGW_SRG_self_energy(500d0, nBas, nBas, nC, nO, nV, nR, nS, eGW, Om, rho, EcGM, SigC, Z)
  else
    call RGW_self_energy(eta, nBas, nBas, nC, nO, nV, nR, nS, eGW, Om, rho, EcGM, SigC, Z)
  end if
  SigC = 0.5d0*(SigC + transpose(SigC))
  call MOtoAO(nBas, nBas, S, c, SigC, SigCp)
  F = T + V + J + 0.5d0*K + SigCp
  cp = matmul(transpose(X), matmul(F, X))
  call diagonalize_matrix(nBas, cp, eNew)
  flush(6)
  deallocate(J, K, ERI_MO, Aph, Bph, Om, XpY, XmY, rho, SigC, SigCp, Z, cp)
end subroutine

! GW/UG0W0.f90:92-150 (nC = nR = 0, no printing), composed of QuAcK's own
! unrestricted routines in the same order.  EcGM: the non-SRG routine returns
! one value per spin; the SRG routine one scalar (their sum), put in EcGM(1).
subroutine qkref_ug0w0(lin, tda_w, srg, eta, nBas, nOa, nOb, nCa, nCb, nRa, nRb, ERI_aaaa, ERI_aabb, ERI_bbbb, &
                       eHF, eGW, SigC, Z, EcRPA, EcGM) bind(C, name='qkref_ug0w0')
  use iso_c_binding
  implicit none
  integer(c_int), value :: lin, tda_w, srg, nBas, nOa, nOb, nCa, nCb, nRa, nRb
  real(c_double), value :: eta
  real(c_double), intent(in) :: ERI_aaaa(nBas,nBas,nBas,nBas), ERI_aabb(nBas,nBas,nBas,nBas)
  real(c_double), intent(in) :: ERI_bbbb(nBas,nBas,nBas,nBas), eHF(nBas,2)
  real(c_double), intent(out) :: eGW(nBas,2), SigC(nBas,2), Z(nBas,2), EcRPA, EcGM(2)
  integer :: nC(2), nO(2), nV(2), nR(2), nS(2), nSa, nSb, nSt, is
  double precision, allocatable :: Aph(:,:), Bph(:,:), Om(:), XpY(:,:), XmY(:,:), rho(:,:,:,:), eGWlin(:,:)
  double precision :: ecs
  logical :: TDAW, doSRG
  TDAW = tda_w /= 0
  doSRG = srg /= 0
  nC(1) = nCa; nC(2) = nCb; nR(1) = nRa; nR(2) = nRb
  nO(1) = nOa; nO(2) = nOb
  nV = nBas - nO
  nS = (nO - nC)*(nV - nR)
  nSa = nS(1); nSb = nS(2); nSt = nSa + nSb
  allocate(Aph(nSt,nSt), Bph(nSt,nSt), Om(nSt), XpY(nSt,nSt), XmY(nSt,nSt), rho(nBas,nBas,nSt,2), eGWlin(nBas,2))
  Bph = 0d0
  call phULR_A(1, .true., nBas, nC, nO, nV, nR, nSa, nSb, nSt, 1d0, eHF, ERI_aaaa, ERI_aabb, ERI_bbbb, Aph)
  if(.not. TDAW) call phULR_B(1, .true., nBas, nC, nO, nV, nR, nSa, nSb, nSt, 1d0, ERI_aaaa, ERI_aabb, ERI_bbbb, Bph)
  call phULR(TDAW, nSa, nSb, nSt, Aph, Bph, EcRPA, Om, XpY, XmY)
  call UGW_excitation_density(nBas, nC, nO, nR, nSa, nSb, nSt, ERI_aaaa, ERI_aabb, ERI_bbbb, XpY, rho)
  EcGM = 0d0
  if(doSRG) then
    call UGW_SRG_self_energy_diag(500d0, nBas, nC, nO, nV, nR, nSt, eHF, Om, rho, ecs, SigC, Z)
    EcGM(1) = ecs
  else
    call UGW_self_energy_diag(eta, nBas, nC, nO, nV, nR, nSt, eHF, Om, rho, EcGM, SigC, Z)
  end if
  eGWlin = eHF + Z*SigC
  if(lin /= 0) then
    eGW = eGWlin
  else
    do is = 1, 2
      call UGW_QP_graph(doSRG, eta, 500d0, nBas, nC(is), nO(is), nV(is), nR(is), nSt, eHF(:,is), Om, &
                        rho(:,:,:,is), eGWlin(:,is), eHF(:,is), eGW(:,is), Z(:,is))
    end do
  end if
  Bph = 0d0
  call phULR_A(1, .true., nBas, nC, nO, nV, nR, nSa, nSb, nSt, 1d0, eGW, ERI_aaaa, ERI_aabb, ERI_bbbb, Aph)
  if(.not. TDAW) call phULR_B(1, .true., nBas, nC, nO, nV, nR, nSa, nSb, nSt, 1d0, ERI_aaaa, ERI_aabb, ERI_bbbb, Bph)
  call phULR(TDAW, nSa, nSb, nSt, Aph, Bph, EcRPA, Om, XpY, XmY)
  flush(6)
  deallocate(Aph, Bph, Om, XpY, XmY, rho, eGWlin)
end subroutine

! GW/UGW_phBSE.f90:60-196 with the static kernel (no dynamical correction, no
! transition vectors), composed of QuAcK's own routines in the same order:
! spin-conserved dRPA screening at eW, its excitation densities, then the
! spin-conserved (ispin 1) and spin-flip (ispin 2) BSE at eGW with dRPA =
! false plus the static kernels, each solved by phULR.
subroutine qkref_uphbse(tda_w, tda, dosc, dosf, eta, nBas, nOa, nOb, nCa, nCb, nRa, nRb, ERI_aaaa, ERI_aabb, &
                        ERI_bbbb, eW, eGW, EcBSE, OmSC, OmSF) bind(C, name='qkref_uphbse')
  use iso_c_binding
  implicit none
  integer(c_int), value :: tda_w, tda, dosc, dosf, nBas, nOa, nOb, nCa, nCb, nRa, nRb
  real(c_double), value :: eta
  real(c_double), intent(in) :: ERI_aaaa(nBas,nBas,nBas,nBas), ERI_aabb(nBas,nBas,nBas,nBas)
  real(c_double), intent(in) :: ERI_bbbb(nBas,nBas,nBas,nBas), eW(nBas,2), eGW(nBas,2)
  real(c_double), intent(out) :: EcBSE(2), OmSC(*), OmSF(*)
  integer :: nC(2), nO(2), nV(2), nR(2), nS(2), nS_aa, nS_bb, nS_sc, nS_ab, nS_ba, nS_sf
  double precision, allocatable :: Aph(:,:), Bph(:,:), KA(:,:), KB(:,:), OmRPA(:), XpY(:,:), XmY(:,:), rho(:,:,:,:)
  double precision, allocatable :: Om(:)
  double precision :: EcRPA
  logical :: TDAW, TDAB
  TDAW = tda_w /= 0
  TDAB = tda /= 0
  nC(1) = nCa; nC(2) = nCb; nR(1) = nRa; nR(2) = nRb
  nO(1) = nOa; nO(2) = nOb
  nV = nBas - nO
  nS = (nO - nC)*(nV - nR)
  nS_aa = nS(1); nS_bb = nS(2); nS_sc = nS_aa + nS_bb
  nS_ab = (nO(1) - nC(1))*(nV(2) - nR(2))
  nS_ba = (nO(2) - nC(2))*(nV(1) - nR(1))
  nS_sf = nS_ab + nS_ba
  EcBSE = 0d0
  allocate(Aph(nS_sc,nS_sc), Bph(nS_sc,nS_sc), OmRPA(nS_sc), XpY(nS_sc,nS_sc), XmY(nS_sc,nS_sc), &
           rho(nBas,nBas,nS_sc,2))
  Bph = 0d0
  call phULR_A(1, .true., nBas, nC, nO, nV, nR, nS_aa, nS_bb, nS_sc, 1d0, eW, ERI_aaaa, ERI_aabb, ERI_bbbb, Aph)
  if(.not. TDAW) call phULR_B(1, .true., nBas, nC, nO, nV, nR, nS_aa, nS_bb, nS_sc, 1d0, ERI_aaaa, ERI_aabb, &
                              ERI_bbbb, Bph)
  call phULR(TDAW, nS_aa, nS_bb, nS_sc, Aph, Bph, EcRPA, OmRPA, XpY, XmY)
  call UGW_excitation_density(nBas, nC, nO, nR, nS_aa, nS_bb, nS_sc, ERI_aaaa, ERI_aabb, ERI_bbbb, XpY, rho)
  deallocate(Aph, Bph, XpY, XmY)
  if(dosc /= 0) then
    allocate(Aph(nS_sc,nS_sc), Bph(nS_sc,nS_sc), KA(nS_sc,nS_sc), KB(nS_sc,nS_sc), Om(nS_sc), XpY(nS_sc,nS_sc), &
             XmY(nS_sc,nS_sc))
    Bph = 0d0; KB = 0d0
    call phULR_A(1, .false., nBas, nC, nO, nV, nR, nS_aa, nS_bb, nS_sc, 1d0, eGW, ERI_aaaa, ERI_aabb, ERI_bbbb, Aph)
    if(.not. TDAB) call phULR_B(1, .false., nBas, nC, nO, nV, nR, nS_aa, nS_bb, nS_sc, 1d0, ERI_aaaa, ERI_aabb, &
                                ERI_bbbb, Bph)
    call UGW_phBSE_static_kernel_A(1, eta, nBas, nC, nO, nV, nR, nS_aa, nS_bb, nS_sc, nS_sc, 1d0, OmRPA, rho, KA)
    if(.not. TDAB) call UGW_phBSE_static_kernel_B(1, eta, nBas, nC, nO, nV, nR, nS_aa, nS_bb, nS_sc, nS_sc, 1d0, &
                                                  OmRPA, rho, KB)
    Aph = Aph + KA
    if(.not. TDAB) Bph = Bph + KB
    call phULR(TDAB, nS_aa, nS_bb, nS_sc, Aph, Bph, EcBSE(1), Om, XpY, XmY)
    OmSC(1:nS_sc) = Om
    deallocate(Aph, Bph, KA, KB, Om, XpY, XmY)
  end if
  if(dosf /= 0) then
    allocate(Aph(nS_sf,nS_sf), Bph(nS_sf,nS_sf), KA(nS_sf,nS_sf), KB(nS_sf,nS_sf), Om(nS_sf), XpY(nS_sf,nS_sf), &
             XmY(nS_sf,nS_sf))
    Bph = 0d0; KB = 0d0
    call phULR_A(2, .false., nBas, nC, nO, nV, nR, nS_ab, nS_ba, nS_sf, 1d0, eGW, ERI_aaaa, ERI_aabb, ERI_bbbb, Aph)
    if(.not. TDAB) call phULR_B(2, .false., nBas, nC, nO, nV, nR, nS_ab, nS_ba, nS_sf, 1d0, ERI_aaaa, ERI_aabb, &
                                ERI_bbbb, Bph)
    call UGW_phBSE_static_kernel_A(2, eta, nBas, nC, nO, nV, nR, nS_ab, nS_ba, nS_sf, nS_sc, 1d0, OmRPA, rho, KA)
    if(.not. TDAB) call UGW_phBSE_static_kernel_B(2, eta, nBas, nC, nO, nV, nR, nS_ab, nS_ba, nS_sf, nS_sc, 1d0, &
                                                  OmRPA, rho, KB)
    Aph = Aph + KA
    if(.not. TDAB) Bph = Bph + KB
    call phULR(TDAB, nS_ab, nS_ab, nS_sf, Aph, Bph, EcBSE(2), Om, XpY, XmY)
    OmSF(1:nS_sf) = Om
    deallocate(Aph, Bph, KA, KB, Om, XpY, XmY)
  end if
  flush(6)
  deallocate(OmRPA, rho)
end subroutine

! evUGW's cycle (GW/evUGW.f90:103-215) composed of QuAcK's own routines, with
! TDA = TDA_W for the B build and the printing left out
subroutine qkref_evugw(maxSCF, thresh, max_diis, lin, tda_w, srg, eta, nBas, nOa, nOb, nCa, nCb, nRa, nRb, &
                       ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, eGW, SigC, Z, EcRPA, EcGM, nSCF_out) &
                       bind(C, name='qkref_evugw')
  use iso_c_binding
  implicit none
  integer(c_int), value :: maxSCF, max_diis, lin, tda_w, srg, nBas, nOa, nOb, nCa, nCb, nRa, nRb
  real(c_double), value :: thresh, eta
  real(c_double), intent(in) :: ERI_aaaa(nBas,nBas,nBas,nBas), ERI_aabb(nBas,nBas,nBas,nBas)
  real(c_double), intent(in) :: ERI_bbbb(nBas,nBas,nBas,nBas), eHF(nBas,2)
  real(c_double), intent(out) :: eGW(nBas,2), SigC(nBas,2), Z(nBas,2), EcRPA, EcGM(2)
  integer(c_int), intent(out) :: nSCF_out
  integer :: nC(2), nO(2), nV(2), nR(2), nS(2), nSa, nSb, nSt, is, nSCF, n_diis
  double precision, allocatable :: Aph(:,:), Bph(:,:), Om(:), XpY(:,:), XmY(:,:), rho(:,:,:,:), eOld(:,:)
  double precision, allocatable :: error_diis(:,:,:), e_diis(:,:,:)
  double precision :: ecs, Conv, rcond(2)
  logical :: TDAW, doSRG
  TDAW = tda_w /= 0
  doSRG = srg /= 0
  nC(1) = nCa; nC(2) = nCb; nR(1) = nRa; nR(2) = nRb
  nO(1) = nOa; nO(2) = nOb
  nV = nBas - nO
  nS = (nO - nC)*(nV - nR)
  nSa = nS(1); nSb = nS(2); nSt = nSa + nSb
  allocate(Aph(nSt,nSt), Bph(nSt,nSt), Om(nSt), XpY(nSt,nSt), XmY(nSt,nSt), rho(nBas,nBas,nSt,2), eOld(nBas,2))
  allocate(error_diis(nBas,max(max_diis,1),2), e_diis(nBas,max(max_diis,1),2))
  error_diis = 0d0; e_diis = 0d0
  nSCF = 0; n_diis = 0; Conv = 1d0
  eGW = eHF; eOld = eHF; Z = 1d0; EcGM = 0d0
  do while(Conv > thresh .and. nSCF <= maxSCF)
    Bph = 0d0
    call phULR_A(1, .true., nBas, nC, nO, nV, nR, nSa, nSb, nSt, 1d0, eGW, ERI_aaaa, ERI_aabb, ERI_bbbb, Aph)
    if(.not. TDAW) call phULR_B(1, .true., nBas, nC, nO, nV, nR, nSa, nSb, nSt, 1d0, ERI_aaaa, ERI_aabb, ERI_bbbb, Bph)
    call phULR(TDAW, nSa, nSb, nSt, Aph, Bph, EcRPA, Om, XpY, XmY)
    call UGW_excitation_density(nBas, nC, nO, nR, nSa, nSb, nSt, ERI_aaaa, ERI_aabb, ERI_bbbb, XpY, rho)
    if(doSRG) then
      call UGW_SRG_self_energy_diag(500d0, nBas, nC, nO, nV, nR, nSt, eGW, Om, rho, ecs, SigC, Z)
      EcGM(1) = ecs
    else
      call UGW_self_energy_diag(eta, nBas, nC, nO, nV, nR, nSt, eGW, Om, rho, EcGM, SigC, Z)
    end if
    if(lin /= 0) then
      eGW = eHF + SigC
    else
      do is = 1, 2
        call UGW_QP_graph(doSRG, eta, 500d0, nBas, nC(is), nO(is), nV(is), nR(is), nSt, eHF(:,is), Om, &
                          rho(:,:,:,is), eOld(:,is), eOld(:,is), eGW(:,is), Z(:,is))
      end do
    end if
    Conv = maxval(abs(eGW - eOld))
    if(max_diis > 1) then
      n_diis = min(n_diis+1, max_diis)
      do is = 1, 2
        call DIIS_extrapolation(rcond(1), nBas, nBas, n_diis, error_diis(:,1:n_diis,is), e_diis(:,1:n_diis,is), &
                                eGW(:,is) - eOld(:,is), eGW(:,is))
      end do
    end if
    eOld = eGW
    nSCF = nSCF + 1
  end do
  nSCF_out = nSCF
  flush(6)
  deallocate(Aph, Bph, Om, XpY, XmY, rho, eOld, error_diis, e_diis)
end subroutine

