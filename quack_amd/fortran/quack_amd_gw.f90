! ISO_C_BINDING interface to libquack_amd.so (include/quack_amd.h).
!
! Same style as QuAcK's own bind(C) module for its GPU library
! (src/GPU/cu_quack_module.f90:11-36): value integers, c_double arrays passed
! by reference, an opaque handle as type(c_ptr).  All arrays keep QuAcK's
! Fortran layout, so no transposes are needed on either side.

module quack_amd_gw

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: QK_OK = 0
  integer(c_int), parameter :: QK_E_ARG = 1
  integer(c_int), parameter :: QK_E_HIP = 2
  integer(c_int), parameter :: QK_E_EIG = 3
  integer(c_int), parameter :: QK_E_NOCONV = 4
  integer(c_int), parameter :: QK_E_UNSUPPORTED = 5
  integer(c_int), parameter :: QK_E_ALLOC = 6
  integer(c_int), parameter :: QK_E_UNSTABLE = 7

  interface

    integer(c_int) function qk_gw_create(h, nBas, nOrb, nC, nO, nV, nR, nS) bind(C, name='qk_gw_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: nBas, nOrb, nC, nO, nV, nR, nS
    end function

    integer(c_int) function qk_gw_destroy(h) bind(C, name='qk_gw_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    type(c_ptr) function qk_gw_last_error(h) bind(C, name='qk_gw_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function qk_gw_set_eri_ao(h, ERI_AO, c) bind(C, name='qk_gw_set_eri_ao')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ERI_AO(*), c(*)
    end function

    integer(c_int) function qk_gw_set_eri_mo(h, ERI_MO) bind(C, name='qk_gw_set_eri_mo')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ERI_MO(*)
    end function

    integer(c_int) function qk_rg0w0(h, eHF, eta, linearize, TDA_W, doSRG, eGW, SigC, Z, Om, EcRPA, EcGM) &
        bind(C, name='qk_rg0w0')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: eHF(*)
      real(c_double), value :: eta
      integer(c_int), value :: linearize, TDA_W, doSRG
      real(c_double), intent(out) :: eGW(*), SigC(*), Z(*), Om(*), EcRPA, EcGM
    end function

    integer(c_int) function qk_evrgw(h, eHF, eGW_start, maxSCF, thresh, max_diis, eta, linearize, TDA_W, doSRG, &
                                     eGW, SigC, Z, Om, EcRPA, EcGM, nSCF) bind(C, name='qk_evrgw')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: eHF(*)
      type(c_ptr), value :: eGW_start
      integer(c_int), value :: maxSCF, max_diis, linearize, TDA_W, doSRG
      real(c_double), value :: thresh, eta
      real(c_double), intent(out) :: eGW(*), SigC(*), Z(*), Om(*), EcRPA, EcGM
      integer(c_int), intent(out) :: nSCF
    end function

    ! static phBSE@GW (GW/RGW_phBSE.f90) from the MO integrals
    integer(c_int) function qk_rgw_phbse(dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet, eta, &
                                         nOrb, nC, nO, nV, nR, nS, ERI, eW, eGW, EcBSE, OmBSE) bind(C, name='qk_rgw_phbse')
      import :: c_int, c_double
      integer(c_int), value :: dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet
      real(c_double), value :: eta
      integer(c_int), value :: nOrb, nC, nO, nV, nR, nS
      real(c_double), intent(in) :: ERI(*), eW(*), eGW(*)
      real(c_double), intent(out) :: EcBSE(2), OmBSE(*)
    end function

    integer(c_int) function qk_rgw_phbse_dyn(dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet, &
                                             eta, nOrb, nC, nO, nV, nR, nS, ERI, eW, eGW, EcBSE, OmBSE, OmDyn, ZDyn) &
                                             bind(C, name='qk_rgw_phbse_dyn')
      import :: c_int, c_double
      integer(c_int), value :: dophBSE2, exchange_kernel, TDA_W, TDA, dBSE, dTDA, singlet, triplet
      real(c_double), value :: eta
      integer(c_int), value :: nOrb, nC, nO, nV, nR, nS
      real(c_double), intent(in) :: ERI(*), eW(*), eGW(*)
      real(c_double), intent(out) :: EcBSE(2), OmBSE(*), OmDyn(20), ZDyn(20)
    end function

    ! unrestricted G0W0 and evGW (GW/UG0W0.f90, GW/evUGW.f90); nC, nO, nR per spin
    integer(c_int) function qk_ug0w0(nBas, nC, nO, nR, ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, eta, linearize, TDA_W, &
                                     doSRG, eGW, SigC, Z, Om, EcRPA, EcGM) bind(C, name='qk_ug0w0')
      import :: c_int, c_double
      integer(c_int), value :: nBas, linearize, TDA_W, doSRG
      integer(c_int), intent(in) :: nC(2), nO(2), nR(2)
      real(c_double), intent(in) :: ERI_aaaa(*), ERI_aabb(*), ERI_bbbb(*), eHF(*)
      real(c_double), value :: eta
      real(c_double), intent(out) :: eGW(*), SigC(*), Z(*), Om(*), EcRPA, EcGM(2)
    end function

    integer(c_int) function qk_evugw(nBas, nC, nO, nR, ERI_aaaa, ERI_aabb, ERI_bbbb, eHF, maxSCF, thresh, max_diis, &
                                     eta, linearize, TDA_W, doSRG, eGW, SigC, Z, Om, EcRPA, EcGM, nSCF) &
                                     bind(C, name='qk_evugw')
      import :: c_int, c_double
      integer(c_int), value :: nBas, maxSCF, max_diis, linearize, TDA_W, doSRG
      integer(c_int), intent(in) :: nC(2), nO(2), nR(2)
      real(c_double), intent(in) :: ERI_aaaa(*), ERI_aabb(*), ERI_bbbb(*), eHF(*)
      real(c_double), value :: thresh, eta
      real(c_double), intent(out) :: eGW(*), SigC(*), Z(*), Om(*), EcRPA, EcGM(2)
      integer(c_int), intent(out) :: nSCF
    end function

    ! unrestricted static phBSE (GW/UGW_phBSE.f90)
    integer(c_int) function qk_ugw_phbse(nBas, nC, nO, nR, ERI_aaaa, ERI_aabb, ERI_bbbb, eW, eGW, eta, TDA_W, TDA, &
                                         spin_conserved, spin_flip, exchange_kernel, EcBSE, OmSC, OmSF) &
                                         bind(C, name='qk_ugw_phbse')
      import :: c_int, c_double
      integer(c_int), value :: nBas, TDA_W, TDA, spin_conserved, spin_flip, exchange_kernel
      integer(c_int), intent(in) :: nC(2), nO(2), nR(2)
      real(c_double), intent(in) :: ERI_aaaa(*), ERI_aabb(*), ERI_bbbb(*), eW(*), eGW(*)
      real(c_double), value :: eta
      real(c_double), intent(out) :: EcBSE(2), OmSC(*), OmSF(*)
    end function

    ! RHF Fock builds (HF/RHF.f90:178-179) on the device
    integer(c_int) function qk_rhf_create(h, nBas) bind(C, name='qk_rhf_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: nBas
    end function

    integer(c_int) function qk_rhf_destroy(h) bind(C, name='qk_rhf_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function qk_rhf_set_eri_ao(h, ERI_AO) bind(C, name='qk_rhf_set_eri_ao')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ERI_AO(*)
    end function

    integer(c_int) function qk_rhf_jk(h, ERI_device, n_eri, P, J, K) bind(C, name='qk_rhf_jk')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h, ERI_device
      integer(c_long_long), value :: n_eri
      real(c_double), intent(in) :: P(*)
      real(c_double), intent(out) :: J(*), K(*)
    end function

    ! full AO->MO transformation of the integrals (AOtoMO/AOtoMO_ERI_RHF.f90)
    integer(c_int) function qk_aotomo_eri_rhf(nBas, nOrb, c, ERI_AO, ERI_MO) bind(C, name='qk_aotomo_eri_rhf')
      import :: c_int, c_double
      integer(c_int), value :: nBas, nOrb
      real(c_double), intent(in) :: c(*), ERI_AO(*)
      real(c_double), intent(out) :: ERI_MO(*)
    end function

    ! quasiparticle self-consistent GW (GW/qsRGW.f90): one handle per calculation,
    ! one call per pass of the loop body
    integer(c_int) function qk_qsgw_create(h, nBas, nOrb, nC, nO, nV, nR, nS, max_diis, S, T, V, Hc, X, ERI_AO, &
                                           PHF, cHF, eHF) bind(C, name='qk_qsgw_create')
      import :: c_int, c_ptr, c_double
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: nBas, nOrb, nC, nO, nV, nR, nS, max_diis
      real(c_double), intent(in) :: S(*), T(*), V(*), Hc(*), X(*), ERI_AO(*), PHF(*), cHF(*), eHF(*)
    end function

    integer(c_int) function qk_qsgw_cycle(h, eta, TDA_W, doSRG, eGW, c, SigC, Z, E, EcRPA, EcGM, err_max) &
                                          bind(C, name='qk_qsgw_cycle')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: eta
      integer(c_int), value :: TDA_W, doSRG
      real(c_double), intent(out) :: eGW(*), c(*), SigC(*), Z(*), E(5), EcRPA, EcGM, err_max
    end function

    integer(c_int) function qk_qsgw_destroy(h) bind(C, name='qk_qsgw_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

  end interface

contains

  ! Stop with a nonzero code and the library's message (QuAcK's CUDA helpers
  ! exit(0) on errors, src/cuda/src/utils.cu:10-16; here the status survives).
  subroutine qk_check(rc, h, where)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if(rc == QK_OK) return
    text = ''
    call c_f_pointer(qk_gw_last_error(h), msg, [512])
    do i = 1, 512
      if(msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*,'(A,A,A,I0,A,A)') ' quack_amd: ', where, ' failed with status ', rc, ': ', trim(text)
    error stop 70
  end subroutine

end module
