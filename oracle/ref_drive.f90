! Test harness over the reference's own Fortran (test infrastructure only: loaded by the
! tests that check the C oracle against the reference, never by the product).
! oracle/build_ref.py compiles the reference's MDSTEDC chain with this file into
! oracle/_ref/libhssref.so.  The entry points take plain C pointers:
!
!   refd_stedc(n, d, e, z, ldz, info)   MDSTEDC('I') (mdstedc.f90:139), tracing every merge
!   refd_trace(cap, n, k, hss, nrot)    the merges traced by the last refd_stedc, in call
!                                       order: order N, deflated size K, 1 for the HSS
!                                       branch, Givens rotations of the deflation
!   refd_hss_build(...)                 the construction half of DHSSEVC
!                                       (CauchyHssEig.f90:2536-2575) on given generators:
!                                       tree, per-node generator dimensions, X := X * H
!
! The trace comes from two link-time hooks (the build renames the symbols in the
! reference's objects): mdlaed1's call of MDLAED2 (mdlaed1.f90:184) goes to refw_mdlaed2,
! and mdlaed2's DROT (mdlaed2.f90:352, one call per deflation rotation) to refw_drot.

module refdrv_trace
    implicit none
    integer, parameter :: tcap = 65536
    integer :: nmerge = 0, rot_count = 0
    integer :: tr_n(tcap), tr_k(tcap), tr_hss(tcap), tr_rot(tcap)
end module refdrv_trace

! MDLAED2 with its outcome recorded.  HFLAG .FALSE. is the HSS branch (mdlaed2.f90:134-137).
subroutine refw_mdlaed2(k, n, n1, d, q, ldq, indxq, rho, z, dlamda, w, q2, indx, indxc, indxp, &
                        coltyp, hflag, info)
    use refdrv_trace
    implicit none
    integer :: k, n, n1, ldq, info
    integer :: indxq(*), indx(*), indxc(*), indxp(*), coltyp(*)
    double precision :: rho, d(*), q(ldq, *), z(*), dlamda(*), w(*), q2(*)
    logical :: hflag
    external :: mdlaed2
    rot_count = 0
    call mdlaed2(k, n, n1, d, q, ldq, indxq, rho, z, dlamda, w, q2, indx, indxc, indxp, coltyp, hflag, info)
    if (nmerge < tcap) then
        nmerge = nmerge + 1
        tr_n(nmerge) = n
        tr_k(nmerge) = k
        tr_rot(nmerge) = rot_count
        if (hflag) then
            tr_hss(nmerge) = 0
        else
            tr_hss(nmerge) = 1
        end if
    end if
end subroutine refw_mdlaed2

subroutine refw_drot(n, x, incx, y, incy, c, s)
    use refdrv_trace
    implicit none
    integer :: n, incx, incy
    double precision :: x(*), y(*), c, s
    external :: drot
    rot_count = rot_count + 1
    call drot(n, x, incx, y, incy, c, s)
end subroutine refw_drot

subroutine refd_stedc(n, d, e, z, ldz, info) bind(c, name="refd_stedc")
    use iso_c_binding
    use refdrv_trace
    implicit none
    integer(c_int), value :: n, ldz
    real(c_double) :: d(*), e(*), z(ldz, *)
    integer(c_int) :: info
    double precision, allocatable :: work(:)
    integer, allocatable :: iwork(:)
    double precision :: wq(1)
    integer :: iq(1), lwork, liwork
    external :: mdstedc
    nmerge = 0
    call mdstedc('I', n, d, e, z, ldz, wq, -1, iq, -1, info)
    if (info /= 0) return
    ! the query returns DSTEDC's 1 + 4N + N^2 (mdstedc.f90:225-227), but an HSS-branch
    ! merge lays DIFL, DIFR, ALPHA and MDLAED3's work out after its K x N Q2
    ! (mdlaed1.f90:205-209), about 2N + K beyond it: give 8N more
    lwork = max(1, int(wq(1))) + 8 * n
    liwork = max(1, iq(1))
    allocate (work(lwork), iwork(liwork))
    call mdstedc('I', n, d, e, z, ldz, work, lwork, iwork, liwork, info)
    deallocate (work, iwork)
end subroutine refd_stedc

function refd_trace(cap, n, k, hss, nrot) bind(c, name="refd_trace") result(m)
    use iso_c_binding
    use refdrv_trace
    implicit none
    integer(c_int), value :: cap
    integer(c_int) :: n(*), k(*), hss(*), nrot(*)
    integer(c_int) :: m
    m = min(nmerge, cap)
    n(1:m) = tr_n(1:m)
    k(1:m) = tr_k(1:m)
    hss(1:m) = tr_hss(1:m)
    nrot(1:m) = tr_rot(1:m)
end function refd_trace

! HSS approximation of A(i,j) = u(i) v(j) / (f(i) - d0(j)) by the reference's
! cauchy2hssEig_omp over its own tree (npart / GetTreeInfo, set up as DHSSEVC does),
! then X := X * H by fastHssmmL_omp.  Outputs (capacity 2*lm - 1 nodes, lm <= k/ni + 1):
! the post-ordered parent array, the (rows, columns) of U, V and B and the D order of
! every node, the number of nodes and the reference's flop and swap counts.
subroutine refd_hss_build(k, ni, f, d0, u, v, difl, difr, tol, nrow, x, ldx, ltr_out, tr_out, usz, vsz, &
                          bsz, dsz, nflops_out, nswap_out, info) bind(c, name="refd_hss_build")
    use iso_c_binding
    use aux_hss
    use CauchyHssEig
    use ConstructHssd
    implicit none
    integer(c_int), value :: k, ni, nrow, ldx
    real(c_double) :: f(*), d0(*), u(*), v(*), difl(*), difr(*), x(ldx, *)
    real(c_double), value :: tol
    integer(c_int) :: ltr_out, tr_out(*), usz(2, *), vsz(2, *), bsz(2, *), dsz(*)
    integer(c_int) :: nflops_out, nswap_out, info
    integer :: ltr, lm, lvl, n1, pnh, nflops, nswap, tnk, i, nb
    type(HTinfo) :: tre
    type(hssmm) :: ph
    integer, allocatable :: tr(:), m(:)
    double precision, allocatable :: dd(:), hh(:), xx(:, :)
    info = 0
    nb = ni
    ltr = 0
    if (mod(k, nb) >= nb / 2) then
        lm = k / nb + 1
    else
        lm = k / nb
    end if
    allocate (m(lm), tr(2 * lm - 1))
    call npart(k, nb, tr, ltr, m, lm)
    lvl = ceiling(log2(ltr + 1)) - 1
    n1 = (ltr + 1) / 2
    allocate (tre%ttr(lvl, n1), tre%lentr(lvl), tre%lfx(ltr), tre%ch(2, ltr))
    call GetTreeInfo(tr, ltr, m, m, tre, lvl)
    allocate (ph%D(ltr), ph%U(2, ltr), ph%V(2, ltr), ph%B(2, ltr), ph%ph(4, ltr), ph%pd(2, ltr))
    call hssexpmm0(ph, ltr)
    tnk = 0
    do i = 1, lm
        tnk = tnk + m(i)**2
    end do
    allocate (hh(10 * nb * k), dd(tnk))
    pnh = 1
    nflops = 0
    nswap = 0
    call cauchy2hssEig_omp(f, d0, u, v, difl, difr, k, tr, ltr, m, lm, ph, hh, dd, tol, tre, lvl, pnh, nflops, nswap)
    if (nrow > 0) then
        allocate (xx(nrow, k))
        xx(:, :) = x(1:nrow, 1:k)
        call fastHssmmL_omp(dd, hh, ph, m, xx, tr, ltr, nrow, k, tre, lvl)
        x(1:nrow, 1:k) = xx(:, :)
        deallocate (xx)
    end if
    ltr_out = ltr
    tr_out(1:ltr) = tr(1:ltr)
    usz(:, 1:ltr) = ph%U(:, 1:ltr)
    vsz(:, 1:ltr) = ph%V(:, 1:ltr)
    bsz(:, 1:ltr) = ph%B(:, 1:ltr)
    dsz(1:ltr) = ph%D(1:ltr)
    nflops_out = nflops
    nswap_out = nswap
    deallocate (tre%ttr, tre%lentr, tre%lfx, tre%ch)
    deallocate (ph%D, ph%U, ph%V, ph%B, ph%ph, ph%pd)
    deallocate (m, tr, hh, dd)
end subroutine refd_hss_build

! The reference's structured RRLU routines themselves (CauchyHssEig.f90:7, 215), for
! step-by-step comparison with the oracle's restatement on the same generators.
subroutine refd_rrlu(rowcol, f, d, a, v, difl, difr, tol, u, w, ff, pl, pu, m, n, rk, pq, nflops, nswap) &
        bind(c, name="refd_rrlu")
    use iso_c_binding
    use CauchyHssEig
    implicit none
    integer(c_int), value :: rowcol, m, n
    real(c_double), value :: tol
    real(c_double) :: f(*), d(*), a(*), v(*), difl(*), difr(*), u(*), w(*), ff(*)
    integer(c_int) :: pl(*), pu(*), pq(*), rk, nflops, nswap
    if (rowcol == 0) then
        call rrluCauchyEigrow(f, d, a, v, difl, difr, tol, u, w, ff, pl, pu, m, n, rk, pq, nflops, nswap)
    else
        call rrluCauchyEigcol(d, f, a, v, difl, difr, tol, u, w, ff, pl, pu, m, n, rk, pq, nflops, nswap)
    end if
end subroutine refd_rrlu
