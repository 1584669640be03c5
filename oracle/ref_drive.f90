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

! ---- bidiagonal SVD (MDBDSDC, mdbdsdc.f90:1) and banded DC (MDSBEVD, mdsbevd.f90:183) ----
! Hooks: mdlasd1's MDLASD2 call (mdlasd1.f90:208) and mdlasd2's DROT (two per deflation
! rotation, mdlasd2.f90:364-365); mdsbed1's MDSBED2 / MDSBED22 calls (mdsbed1.f90:205,
! 262), whose deflation rotates Q's columns (length N) and Z's rows (shorter): only the
! former are counted.

subroutine refw_mdlasd2(nl, nr, sqre, k, d, z, alpha, beta, u, ldu, vt, ldvt, dsigma, u2, ldu2, vt2, ldvt2, &
                        idxp, idx, idxc, idxq, coltyp, hflag, info)
    use refdrv_trace
    implicit none
    integer :: nl, nr, sqre, k, ldu, ldvt, ldu2, ldvt2, info
    double precision :: alpha, beta
    logical :: hflag
    integer :: idxp(*), idx(*), idxc(*), idxq(*), coltyp(*)
    double precision :: d(*), z(*), u(*), vt(*), dsigma(*), u2(*), vt2(*)
    external :: mdlasd2
    rot_count = 0
    call mdlasd2(nl, nr, sqre, k, d, z, alpha, beta, u, ldu, vt, ldvt, dsigma, u2, ldu2, vt2, ldvt2, &
                 idxp, idx, idxc, idxq, coltyp, hflag, info)
    if (nmerge < tcap) then
        nmerge = nmerge + 1
        tr_n(nmerge) = nl + nr + 1
        tr_k(nmerge) = k
        tr_rot(nmerge) = rot_count / 2
        if (hflag) then
            tr_hss(nmerge) = 0
        else
            tr_hss(nmerge) = 1
        end if
    end if
end subroutine refw_mdlasd2

subroutine refw_drot_all(n, x, incx, y, incy, c, s)
    use refdrv_trace
    implicit none
    integer :: n, incx, incy
    double precision :: x(*), y(*), c, s
    external :: drot
    rot_count = rot_count + 1
    call drot(n, x, incx, y, incy, c, s)
end subroutine refw_drot_all

module refdrv_band
    implicit none
    integer :: cur_n = -1
end module refdrv_band

subroutine refw_drot_q(n, x, incx, y, incy, c, s)
    use refdrv_trace
    use refdrv_band
    implicit none
    integer :: n, incx, incy
    double precision :: x(*), y(*), c, s
    external :: drot
    if (n == cur_n) rot_count = rot_count + 1
    call drot(n, x, incx, y, incy, c, s)
end subroutine refw_drot_q

subroutine refw_band_record(n, k, hss)
    use refdrv_trace
    implicit none
    integer :: n, k, hss
    if (nmerge < tcap) then
        nmerge = nmerge + 1
        tr_n(nmerge) = n
        tr_k(nmerge) = k
        tr_rot(nmerge) = rot_count
        tr_hss(nmerge) = hss
    end if
end subroutine refw_band_record

subroutine refw_mdsbed2(k, n, n1, kd, d, q, ldq, indxq, rho, z, z2, dlamda, w, q2, indx, indxc, indxp, &
                        coltyp, hflag, bsmlz, info)
    use refdrv_trace
    use refdrv_band
    implicit none
    integer :: k, n, n1, kd, ldq, bsmlz, info
    double precision :: rho
    logical :: hflag
    integer :: indxq(*), indx(*), indxc(*), indxp(*), coltyp(*)
    double precision :: d(*), q(*), z(*), z2(*), dlamda(*), w(*), q2(*)
    external :: mdsbed2, refw_band_record
    rot_count = 0
    cur_n = n
    call mdsbed2(k, n, n1, kd, d, q, ldq, indxq, rho, z, z2, dlamda, w, q2, indx, indxc, indxp, coltyp, &
                 hflag, bsmlz, info)
    cur_n = -1
    if (hflag) then
        call refw_band_record(n, k, 0)
    else
        call refw_band_record(n, k, 1)
    end if
end subroutine refw_mdsbed2

subroutine refw_mdsbed22(k, n, n1, sb, d, q, ldq, rho, z, z2, dlamda, w, q2, indxq, indxp, info)
    use refdrv_trace
    use refdrv_band
    implicit none
    integer :: k, n, n1, sb, ldq, info
    double precision :: rho
    integer :: indxq(*), indxp(*)
    double precision :: d(*), q(*), z(*), z2(*), dlamda(*), w(*), q2(*)
    external :: mdsbed22, refw_band_record
    rot_count = 0
    cur_n = n
    call mdsbed22(k, n, n1, sb, d, q, ldq, rho, z, z2, dlamda, w, q2, indxq, indxp, info)
    cur_n = -1
    call refw_band_record(n, k, -1)
end subroutine refw_mdsbed22

subroutine refd_bdsdc(n, d, e, u, vt, info) bind(c, name="refd_bdsdc")
    use iso_c_binding
    use refdrv_trace
    implicit none
    integer(c_int), value :: n
    real(c_double) :: d(*), e(*), u(n, *), vt(n, *)
    integer(c_int) :: info
    double precision, allocatable :: work(:)
    integer, allocatable :: iwork(:)
    double precision :: q(1)
    integer :: iq(1)
    external :: mdbdsdc
    nmerge = 0
    ! 3N^2 + 4N for COMPQ = 'I' (mdbdsdc.f90:113), doubled for the HSS branch's generators
    allocate (work(6 * n * n + 16 * n + 16), iwork(8 * n + 16))
    call mdbdsdc('U', 'I', n, d, e, u, n, vt, n, q, iq, work, iwork, info)
    deallocate (work, iwork)
end subroutine refd_bdsdc

subroutine refd_sbevd(n, kd, ab, ldab, w, q, ldq, info) bind(c, name="refd_sbevd")
    use iso_c_binding
    use refdrv_trace
    implicit none
    integer(c_int), value :: n, kd, ldab, ldq
    real(c_double) :: ab(ldab, *), w(*), q(ldq, *)
    integer(c_int) :: info
    double precision, allocatable :: work(:)
    integer, allocatable :: iwork(:)
    double precision :: wq(1)
    integer :: iq(1), lwork, liwork
    external :: mdsbevd
    nmerge = 0
    call mdsbevd('I', 'U', n, kd, ab, ldab, w, q, ldq, wq, -1, iq, -1, info)
    if (info /= 0) return
    lwork = 4 * max(1, int(wq(1))) + (64 + 8 * kd) * n + 16 * kd * kd * n
    liwork = 4 * max(1, iq(1)) + 16 * n
    allocate (work(lwork), iwork(liwork))
    call mdsbevd('I', 'U', n, kd, ab, ldab, w, q, ldq, work, lwork, iwork, liwork, info)
    deallocate (work, iwork)
end subroutine refd_sbevd

! ---- dense -> HSS (mat2hssvd, ConstructHssd.f90:371) and the general Cauchy-like
! construction (cauchy2hss, CauchyHss_VP2.f90:387), set up over npart's tree as DHSSMM and
! CauchyHssMM do; per-node generator dimensions out.  And the two drivers themselves
! (DHSSMM.f90:1, cauchyhssmm.f90:1) for their products.

subroutine refd_tree_dims(ph, ltr, tr, ltr_out, tr_out, usz, vsz, bsz, dsz)
    use aux_hss
    implicit none
    type(hssmm) :: ph
    integer :: ltr, tr(*), ltr_out, tr_out(*), usz(2, *), vsz(2, *), bsz(2, *), dsz(*)
    ltr_out = ltr
    tr_out(1:ltr) = tr(1:ltr)
    usz(:, 1:ltr) = ph%U(:, 1:ltr)
    vsz(:, 1:ltr) = ph%V(:, 1:ltr)
    bsz(:, 1:ltr) = ph%B(:, 1:ltr)
    dsz(1:ltr) = ph%D(1:ltr)
end subroutine refd_tree_dims

subroutine refd_mat2hss(n, a, lda, ni, tol, ltr_out, tr_out, usz, vsz, bsz, dsz, prank, info) &
        bind(c, name="refd_mat2hss")
    use iso_c_binding
    use aux_hss
    use ConstructHssd
    implicit none
    integer(c_int), value :: n, lda, ni
    real(c_double), value :: tol
    real(c_double) :: a(lda, *)
    integer(c_int) :: ltr_out, tr_out(*), usz(2, *), vsz(2, *), bsz(2, *), dsz(*), prank, info
    integer :: ltr, lm, lvl, n1, pnh, tnk, i, nb, rk
    type(HTinfo) :: tre
    type(hssmm) :: ph
    integer, allocatable :: tr(:), blk(:)
    double precision, allocatable :: h(:), dd(:), work(:), acopy(:, :)
    info = 0
    nb = ni
    allocate (blk(n / nb + 1), tr(2 * (n / nb + 1) - 1))
    blk = 0
    ltr = 0
    call npart(n, nb, tr, ltr, blk, lm)
    lvl = ceiling(log2(ltr + 1)) - 1
    n1 = (ltr + 1) / 2
    allocate (tre%ttr(lvl, n1), tre%lentr(lvl), tre%lfx(ltr), tre%ch(2, ltr))
    call GetTreeInfo(tr, ltr, blk, blk, tre, lvl)
    tnk = 0
    do i = 1, lm
        tnk = tnk + blk(i)**2
    end do
    allocate (ph%D(ltr), ph%U(2, ltr), ph%V(2, ltr), ph%B(2, ltr), ph%ph(4, ltr), ph%pd(2, ltr))
    allocate (h(8 * nb * n), dd(tnk), work(2 * lvl * nb * n))
    call hssexpmm0(ph, ltr)
    allocate (acopy(n, n))
    acopy(:, :) = a(1:n, 1:n)
    pnh = 1
    rk = 0
    call mat2hssvd(acopy, n, tr, ltr, blk, lm, ph, h, dd, tol, rk, work, pnh)
    prank = rk
    call refd_tree_dims(ph, ltr, tr, ltr_out, tr_out, usz, vsz, bsz, dsz)
    deallocate (tre%ttr, tre%lentr, tre%lfx, tre%ch, ph%D, ph%U, ph%V, ph%B, ph%ph, ph%pd)
    deallocate (blk, tr, h, dd, work, acopy)
end subroutine refd_mat2hss

subroutine refd_cauchy2hss(n, ni, d, f, u, v, tol, ltr_out, tr_out, usz, vsz, bsz, dsz, info) &
        bind(c, name="refd_cauchy2hss")
    use iso_c_binding
    use aux_hss
    use CauchyHss_VP2
    implicit none
    integer(c_int), value :: n, ni
    real(c_double), value :: tol
    real(c_double) :: d(*), f(*), u(*), v(*)
    integer(c_int) :: ltr_out, tr_out(*), usz(2, *), vsz(2, *), bsz(2, *), dsz(*), info
    integer :: ltr, lm, lvl, n1, pnh, tnk, i, nb, n0
    type(HTinfo) :: tre
    type(hssmm) :: ph
    integer, allocatable :: tr(:), m(:)
    double precision, allocatable :: h(:), dd(:), dc(:), fc(:), uc(:), vc(:)
    info = 0
    nb = ni
    allocate (m(n / nb + 1), tr(2 * (n / nb + 1) - 1))
    m = 0
    ltr = 0
    call npart(n, nb, tr, ltr, m, lm)
    lvl = ceiling(log2(ltr + 1)) - 1
    n1 = (ltr + 1) / 2
    allocate (tre%ttr(lvl, n1), tre%lentr(lvl), tre%lfx(ltr), tre%ch(2, ltr))
    call GetTreeInfo(tr, ltr, m, m, tre, lvl)
    tnk = 0
    do i = 1, lm
        tnk = tnk + m(i)**2
    end do
    allocate (ph%D(ltr), ph%U(2, ltr), ph%V(2, ltr), ph%B(2, ltr), ph%ph(4, ltr), ph%pd(2, ltr))
    allocate (h(8 * nb * n), dd(tnk))
    call hssexpmm0(ph, ltr)
    allocate (dc(n), fc(n), uc(n), vc(n))
    dc = d(1:n)
    fc = f(1:n)
    uc = u(1:n)
    vc = v(1:n)
    pnh = 1
    n0 = n
    call cauchy2hss(dc, fc, uc, vc, n0, tr, ltr, m, lm, ph, h, dd, tol, lvl, pnh)
    call refd_tree_dims(ph, ltr, tr, ltr_out, tr_out, usz, vsz, bsz, dsz)
    deallocate (tre%ttr, tre%lentr, tre%lfx, tre%ch, ph%D, ph%U, ph%V, ph%B, ph%ph, ph%pd)
    deallocate (m, tr, h, dd, dc, fc, uc, vc)
end subroutine refd_cauchy2hss

subroutine refd_dhssmm(n, a, lda, b, ldb, k, tol, ni, ncol, info) bind(c, name="refd_dhssmm")
    use iso_c_binding
    implicit none
    integer(c_int), value :: n, lda, ldb, k, ni, ncol
    real(c_double), value :: tol
    real(c_double) :: a(lda, *), b(ldb, *)
    integer(c_int) :: info
    integer :: prank, nb, nc
    double precision :: time(2)
    external :: dhssmm
    prank = 0
    nb = ni
    nc = ncol
    call dhssmm(a, lda, n, n, b, ldb, k, tol, prank, nb, nc, time, info)
end subroutine refd_dhssmm

subroutine refd_cauchyhssmm(n, d, f, u, v, ni, tol, ncol, b, ldb) bind(c, name="refd_cauchyhssmm")
    use iso_c_binding
    implicit none
    integer(c_int), value :: n, ni, ncol, ldb
    real(c_double), value :: tol
    real(c_double) :: d(*), f(*), u(*), v(*), b(ldb, *)
    integer :: nn, nb, nc
    double precision :: time(2)
    external :: cauchyhssmm
    nn = n
    nb = ni
    nc = ncol
    call cauchyhssmm(d, f, u, v, nn, nb, tol, nc, b, ldb, time)
end subroutine refd_cauchyhssmm
