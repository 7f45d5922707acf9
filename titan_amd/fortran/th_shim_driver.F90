!-------------------------------------------------------------------------------
! th_shim_driver -- exercises mod_titan_hip the way TITAN's calculate_chi does:
! set-up from TITAN-shaped arrays, one eintshechi-equivalent call and the
! Dyson step, then eintshe (field along x, HF prefactor) and the ground-state
! expectation values.  Reads a little-endian stream file (argument 1) and
! writes chiorb_hf, chiorb, cond, tchiorbiikl, rho, mx, my, mz, mp (argument 2).
!-------------------------------------------------------------------------------
program th_shim_driver
  use, intrinsic :: iso_c_binding
  use mod_titan_hip
  implicit none
  integer(c_int) :: hdr(8)
  integer :: nA, nOrb, nhop, n, nk, nx, D, dH, prec
  real(c_double), allocatable :: R(:,:), kimag(:,:), yimag(:), wimag(:), kreal(:,:), wreal(:), x2(:), p2(:)
  integer(c_int), allocatable :: jb(:), ib(:)
  complex(c_double_complex), allocatable :: t(:,:,:), h0(:,:), U(:,:), chi0(:,:), chi(:,:), tchi(:,:), mp(:,:)
  real(c_double), allocatable :: rho(:,:), mx(:,:), my(:,:), mz(:,:)
  real(c_double) :: ue(3)
  real(c_double) :: q(3), sc(4), cond
  character(len=1024) :: fin, fout
  integer :: u1

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u1, file=trim(fin), access="stream", form="unformatted", status="old")
  read(u1) hdr
  nA = hdr(1); nOrb = hdr(2); nhop = hdr(3); n = hdr(4); nk = hdr(5); nx = hdr(6); prec = hdr(7)
  dH = 2*nOrb*nA
  D = 4*nOrb*nOrb*nA
  allocate(R(3,nhop), jb(nhop), ib(nhop), t(nOrb,nOrb,nhop), h0(dH,dH))
  allocate(kimag(3,n), yimag(n), wimag(n), kreal(3,nk), wreal(nk), x2(nx), p2(nx), U(D,D), chi0(D,D), chi(D,D))
  read(u1) R, jb, ib, t, h0, kimag, yimag, wimag, kreal, wreal, q, sc, x2, p2, U
  close(u1)

  call th_setup_from_arrays(nA, nOrb, nhop, R, jb, ib, t, h0, n, kimag, yimag, wimag, nk, kreal, wreal)
  th_precision = int(prec, c_int)
  ! sc = (Ef, eta, etap, omega)
  call th_eintshechi(q, sc(4), sc(1), sc(2), sc(3), x2, p2, chi0)
  call th_rpa_dyson(chi0, U, chi, cond)
  allocate(tchi(D,4), rho(nOrb,nA), mx(nOrb,nA), my(nOrb,nA), mz(nOrb,nA), mp(nOrb,nA))
  ue = [1.0_c_double, 0.0_c_double, 0.0_c_double]
  call th_eintshe(q, sc(4), sc(1), sc(2), sc(3), x2, p2, ue, tchi)
  call th_expectation(sc(1), sc(2), rho, mp, mx, my, mz)
  call th_release()

  open(newunit=u1, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(u1) chi0, chi, cond, tchi, rho, mx, my, mz, mp
  close(u1)
  write(*, "('th_shim_driver: D = ', i0, ', cond = ', es12.5)") D, cond
end program th_shim_driver
