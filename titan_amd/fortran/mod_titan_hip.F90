!-------------------------------------------------------------------------------
! mod_titan_hip -- ISO_C_BINDING interface of libtitan_hip.so (include/titan_hip.h)
! and Fortran wrappers with the argument shapes of TITAN's own arrays.
!
! Drop-in use inside TITAN (see the integration notes of this repository):
!   - after hamilt_local / initConversionMatrices:  call th_setup_from_arrays(...)
!     with h0, the hop table of s%Neighbors(:)%t0i and the (E,k) work list;
!   - the body of eintshechi(q,e) (source/eintshechi.F90:2-178) becomes
!       call generate_real_epoints(e)
!       call th_eintshechi(q, e, s%Ef, eta, etap, x2, p2, chiorb_hf)
!   - the inline Dyson step (source/calculate_chi.F90:84-89) becomes
!       call th_rpa_dyson(chiorb_hf, Umatorb, chiorb)
! Any nonzero status aborts like invers does (source/mod_tools.F90:814-827).
!-------------------------------------------------------------------------------
module mod_titan_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: TH_PREC_FP64 = 0, TH_PREC_MIXED = 1

  type(c_ptr), save :: handle = c_null_ptr
  integer, save :: dimens_ = 0
  integer(c_int), save, public :: th_precision = TH_PREC_FP64

  interface
    integer(c_int) function th_create(nA, nOrb, nhop, R, jb, ib, t, h0, out) bind(C, name="th_create")
      import :: c_int, c_double, c_double_complex, c_ptr
      integer(c_int), value :: nA, nOrb, nhop
      real(c_double), intent(in) :: R(*)
      integer(c_int), intent(in) :: jb(*), ib(*)
      complex(c_double_complex), intent(in) :: t(*), h0(*)
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function th_set_imag_points(h, n, k, y, w) bind(C, name="th_set_imag_points")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_long), value :: n
      real(c_double), intent(in) :: k(*), y(*), w(*)
    end function
    integer(c_int) function th_set_real_kmesh(h, nk, k, w) bind(C, name="th_set_real_kmesh")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_long), value :: nk
      real(c_double), intent(in) :: k(*), w(*)
    end function
    integer(c_int) function th_chi0(h, q, omega, Ef, eta, etap, n_x, x2, p2, precision, chi0) bind(C, name="th_chi0")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: q(3)
      real(c_double), value :: omega, Ef, eta, etap
      integer(c_int), value :: n_x, precision
      real(c_double), intent(in) :: x2(*), p2(*)
      complex(c_double_complex), intent(out) :: chi0(*)
    end function
    integer(c_int) function th_dyson(h, chi0, U, chi, cond) bind(C, name="th_dyson")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: chi0(*), U(*)
      complex(c_double_complex), intent(out) :: chi(*)
      real(c_double), intent(out) :: cond
    end function
    ! traces, cross responses and local-frame rotation (calculate_chi.F90:91-185);
    ! pass c_null_ptr for chi (the last Dyson result), lvec / angles (not
    ! wanted) or any output, c_loc(array) otherwise
    integer(c_int) function th_chi_traces(h, chi, lvec, angles, schi, ls, sl, ll) bind(C, name="th_chi_traces")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, chi, lvec, angles, schi, ls, sl, ll
    end function
    ! <S^{mu nu}_ii> of the ground state (calcSmunu, mod_sumrule.F90:162-245)
    integer(c_int) function th_spin_density(h, Ef, eta, S) bind(C, name="th_spin_density")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      real(c_double), value :: Ef, eta
      complex(c_double_complex), intent(out) :: S(*)
    end function
    ! occupations and magnetisation per orbital (expectation_values_greenfunction,
    ! mod_expectation.F90:28-149)
    integer(c_int) function th_expectation_values(h, Ef, eta, rho, mx, my, mz, mp) &
        bind(C, name="th_expectation_values")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      real(c_double), value :: Ef, eta
      real(c_double), intent(out) :: rho(*), mx(*), my(*), mz(*)
      complex(c_double_complex), intent(out) :: mp(*)
    end function
    ! spin-current response to a field along ue (eintshe, eintshe.F90:2-224);
    ! prefactor: c_null_ptr for the identity, else c_loc of a D x D array
    integer(c_int) function th_current_response(h, q, omega, Ef, eta, etap, n_x, x2, p2, ue, prefactor, tchi) &
        bind(C, name="th_current_response")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: q(3)
      real(c_double), value :: omega, Ef, eta, etap
      integer(c_int), value :: n_x
      real(c_double), intent(in) :: x2(*), p2(*), ue(3)
      type(c_ptr), value :: prefactor
      complex(c_double_complex), intent(out) :: tchi(*)
    end function
    integer(c_int) function th_destroy(h) bind(C, name="th_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function th_last_error() bind(C, name="th_last_error")
      import :: c_ptr
    end function
  end interface

  public :: th_setup_from_arrays, th_eintshechi, th_rpa_dyson, th_eintshe, th_expectation, th_release, &
            th_error_text

contains

  function th_error_text() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: n
    call c_f_pointer(th_last_error(), cs, [4096])
    n = 0
    do while (n < 4096)
      if (cs(n+1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, "('[', a, '] titan_hip status ', i0, ': ', a)") where, rc, th_error_text()
      error stop 1
    end if
  end subroutine

  !> Uploads the model and the (E,k) work list.  Arrays as TITAN holds them:
  !> h0(dimH,dimH), t(nOrb,nOrb,nhop) (Neighbors(n)%t0i(:,:,i) of each hopping
  !> pair), R(3,nhop) (CellVector), jb/ib 1-based atoms (BasisIndex, i);
  !> kimag(3,n), yimag(n), wimag(n) the flattened E_k_imag_mesh with
  !> y(E) and wght(E)*w_k; kreal(3,nk), wreal(nk) the realBZ mesh.
  subroutine th_setup_from_arrays(nA, nOrb, nhop, R, jb, ib, t, h0, n, kimag, yimag, wimag, nk, kreal, wreal)
    integer, intent(in) :: nA, nOrb, nhop, n, nk
    real(c_double), intent(in) :: R(3,nhop), kimag(3,n), yimag(n), wimag(n), kreal(3,nk), wreal(nk)
    integer, intent(in) :: jb(nhop), ib(nhop)
    complex(c_double_complex), intent(in) :: t(nOrb,nOrb,nhop), h0(2*nOrb*nA,2*nOrb*nA)
    integer(c_int), allocatable :: jb0(:), ib0(:)
    if (c_associated(handle)) call th_release()
    allocate(jb0(nhop), ib0(nhop))
    jb0 = int(jb - 1, c_int)
    ib0 = int(ib - 1, c_int)
    call check(th_create(int(nA,c_int), int(nOrb,c_int), int(nhop,c_int), R, jb0, ib0, t, h0, handle), "th_create")
    call check(th_set_imag_points(handle, int(n,c_long), kimag, yimag, wimag), "th_set_imag_points")
    call check(th_set_real_kmesh(handle, int(nk,c_long), kreal, wreal), "th_set_real_kmesh")
    dimens_ = 4*nOrb*nOrb*nA
  end subroutine

  !> chiorb_hf(dimens,dimens) for one frequency e: the eintshechi(q,e) body.
  subroutine th_eintshechi(q, e, Ef, eta, etap, x2, p2, chiorb_hf)
    real(c_double), intent(in) :: q(3), e, Ef, eta, etap, x2(:), p2(:)
    complex(c_double_complex), intent(out) :: chiorb_hf(:,:)
    call check(th_chi0(handle, q, e, Ef, eta, etap, int(size(x2),c_int), x2, p2, th_precision, chiorb_hf), &
               "th_eintshechi")
  end subroutine

  !> chiorb = (identt + chiorb_hf*Umatorb)^-1 * chiorb_hf (calculate_chi.F90:84-89).
  subroutine th_rpa_dyson(chiorb_hf, Umatorb, chiorb, cond)
    complex(c_double_complex), intent(in) :: chiorb_hf(:,:), Umatorb(:,:)
    complex(c_double_complex), intent(out) :: chiorb(:,:)
    real(c_double), intent(out), optional :: cond
    real(c_double) :: c
    call check(th_dyson(handle, chiorb_hf, Umatorb, chiorb, c), "th_rpa_dyson")
    if (present(cond)) cond = c
  end subroutine

  !> tchiorbiikl(dimens,4) for one frequency e: the eintshe(q,e) body for the
  !> field direction ue (ElectricFieldVector); prefactor as calculate_all sets it
  !> (omitted: identity, the HF response).
  subroutine th_eintshe(q, e, Ef, eta, etap, x2, p2, ue, tchiorbiikl, prefactor)
    real(c_double), intent(in) :: q(3), e, Ef, eta, etap, x2(:), p2(:), ue(3)
    complex(c_double_complex), intent(out) :: tchiorbiikl(:,:)
    complex(c_double_complex), intent(in), target, optional :: prefactor(:,:)
    type(c_ptr) :: pp
    pp = c_null_ptr
    if (present(prefactor)) pp = c_loc(prefactor)
    call check(th_current_response(handle, q, e, Ef, eta, etap, int(size(x2),c_int), x2, p2, ue, pp, tchiorbiikl), &
               "th_eintshe")
  end subroutine

  !> rho, mx, my, mz (nOrb,nAtoms) and mp (nOrb,nAtoms) of
  !> expectation_values_greenfunction over the handle's imaginary list.
  subroutine th_expectation(Ef, eta, rho, mp, mx, my, mz)
    real(c_double), intent(in) :: Ef, eta
    real(c_double), intent(out) :: rho(:,:), mx(:,:), my(:,:), mz(:,:)
    complex(c_double_complex), intent(out) :: mp(:,:)
    call check(th_expectation_values(handle, Ef, eta, rho, mx, my, mz, mp), "th_expectation")
  end subroutine

  subroutine th_release()
    integer(c_int) :: rc
    if (c_associated(handle)) rc = th_destroy(handle)
    handle = c_null_ptr
  end subroutine

end module mod_titan_hip
