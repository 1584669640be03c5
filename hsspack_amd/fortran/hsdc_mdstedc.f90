! Fortran drop-ins for HSSPACK's tridiagonal divide-and-conquer entry points
! (SRC/mdstedc.f90, mdlaed0.f90, mdlaed1.f90, mdlaed2.f90, mdlaed3.f90) and its
! bidiagonal SVD driver (SRC/mdbdsdc.f90): the same
! names, argument lists, INFO codes and XERBLA calls, with the work done by the
! MI355X library libhsdc.so through its C interface (include/hsdc.h).
!
! Each call runs on the calling thread's own library context (hsdc_thread_ctx, reset
! to the reference's constants at every call), whose device buffers stay allocated
! between calls, so a repeated call of the same order allocates no device memory;
! hsdc_thread_ctx_release frees them.  WORK/IWORK are accepted for interface
! compatibility (MDSTEDC checks LWORK/LIWORK and answers queries as the original);
! the device needs no host workspace.  XERBLA is the caller's (LAPACK's).

module hsdc_c
   use iso_c_binding
   implicit none
   interface
      integer(c_int) function hsdc_ctx_create(ctx) bind(C, name='hsdc_ctx_create')
         import :: c_int, c_ptr
         type(c_ptr) :: ctx
      end function
      integer(c_int) function hsdc_ctx_destroy(ctx) bind(C, name='hsdc_ctx_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function hsdc_stedc(ctx, compz, n, d, e, z, ldz, work, lwork, iwork, liwork) &
            bind(C, name='hsdc_stedc')
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value :: ctx
         character(kind=c_char), value :: compz
         integer(c_int), value :: n, ldz, lwork, liwork
         real(c_double) :: d(*), e(*), z(*), work(*)
         integer(c_int) :: iwork(*)
      end function
      integer(c_int) function hsdc_laed0(ctx, icompq, qsiz, n, d, e, q, ldq) bind(C, name='hsdc_laed0')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: icompq, qsiz, n, ldq
         real(c_double) :: d(*), e(*), q(*)
      end function
      integer(c_int) function hsdc_deflate(ctx, n, n1, d, z, indxq, rho, hss_min_k, k, rho_out, dlamda, w, &
            indxp, indx, indxc, coltyp, nrot, rot_pj, rot_nj, rot_c, rot_s) bind(C, name='hsdc_deflate')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n, n1, hss_min_k
         real(c_double), value :: rho
         real(c_double) :: d(*), z(*), rho_out, dlamda(*), w(*), rot_c(*), rot_s(*)
         integer(c_int) :: indxq(*), k, indxp(*), indx(*), indxc(*), coltyp(*), nrot, rot_pj(*), rot_nj(*)
      end function
      integer(c_int) function hsdc_laed3(ctx, k, n, n1, d, q, ldq, rho, dlamda, q2, indx, ctot, w) &
            bind(C, name='hsdc_laed3')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: k, n, n1, ldq
         real(c_double), value :: rho
         real(c_double) :: d(*), q(*), dlamda(*), q2(*), w(*)
         integer(c_int) :: indx(*), ctot(*)
      end function
      integer(c_int) function hsdc_loewner(ctx, k, rho, dlamda, z, d, difl, difr, alpha) &
            bind(C, name='hsdc_loewner')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: k
         real(c_double), value :: rho
         real(c_double) :: dlamda(*), z(*), d(*), difl(*), difr(*), alpha(*)
      end function
      integer(c_int) function hsdc_secular(ctx, k, dlamda, w, rho, lam, org, tau, niter) &
            bind(C, name='hsdc_secular')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: k
         real(c_double), value :: rho
         real(c_double) :: dlamda(*), w(*), lam(*), tau(*)
         integer(c_int) :: org(*), niter(*)
      end function
      integer(c_int) function hsdc_hss_evc(ctx, k, ni, tol, f, d, u, v, difl, difr, q, ldq, nrow, ranks) &
            bind(C, name='hsdc_hss_evc')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: k, ni, ldq, nrow
         real(c_double), value :: tol
         real(c_double) :: f(*), d(*), u(*), v(*), difl(*), difr(*), q(*)
         type(c_ptr), value :: ranks
      end function
      type(c_ptr) function hsdc_thread_ctx() bind(C, name='hsdc_thread_ctx')
         import :: c_ptr
      end function
      type(c_ptr) function hsdc_sb_thread_ctx() bind(C, name='hsdc_sb_thread_ctx')
         import :: c_ptr
      end function
      integer(c_int) function hsdc_sbevd(ctx, compz, uplo, n, kd, ab, ldab, d, q, ldq, work, lwork, iwork, &
                                         liwork) bind(C, name='hsdc_sbevd')
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value :: ctx
         character(kind=c_char), value :: compz, uplo
         integer(c_int), value :: n, kd, ldab, ldq, lwork, liwork
         real(c_double) :: ab(*), d(*), q(*), work(*)
         integer(c_int) :: iwork(*)
      end function
      integer(c_int) function hsdc_dhssmm(ctx, a, lda, m, n, b, ldb, k, tol, prank, ni, ncol, time, ranks) &
            bind(C, name='hsdc_dhssmm')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: lda, m, n, ldb, k, prank, ni, ncol
         real(c_double), value :: tol
         real(c_double) :: a(*), b(*), time(*)
         type(c_ptr), value :: ranks
      end function
      integer(c_int) function hsdc_dhssmml(ctx, a, lda, m, n, b, ldb, k, tol, prank, ni, ncol, time, ranks) &
            bind(C, name='hsdc_dhssmml')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: lda, m, n, ldb, k, prank, ni, ncol
         real(c_double), value :: tol
         real(c_double) :: a(*), b(*), time(*)
         type(c_ptr), value :: ranks
      end function
      type(c_ptr) function hsdc_bd_thread_ctx() bind(C, name='hsdc_bd_thread_ctx')
         import :: c_ptr
      end function
      integer(c_int) function hsdc_set_params(ctx, p) bind(C, name='hsdc_set_params')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, p
      end function
      integer(c_int) function hsdc_cauchyhssmm(ctx, d, f, u, v, n, ni, tol, ncol, b, ldb, time, ranks) &
            bind(C, name='hsdc_cauchyhssmm')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: d(*), f(*), u(*), v(*), b(*), time(*)
         integer(c_int), value :: n, ni, ncol, ldb
         real(c_double), value :: tol
         type(c_ptr), value :: ranks
      end function
      integer(c_int) function hsdc_bdsdc(ctx, uplo, compq, n, d, e, u, ldu, vt, ldvt, q, iq, work, iwork) &
            bind(C, name='hsdc_bdsdc')
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value :: ctx
         character(kind=c_char), value :: uplo, compq
         integer(c_int), value :: n, ldu, ldvt
         real(c_double) :: d(*), e(*), u(*), vt(*), q(*), work(*)
         integer(c_int) :: iq(*), iwork(*)
      end function
   end interface

   ! HSSPACK's compiled-in constants: BSMLZ (mdlaed2.f90:175), Ni (mdlaed3.f90:253),
   ! TOL (CauchyHssEig.f90:2523)
   integer, parameter :: hsdc_bsmlz = 1500, hsdc_ni = 200
   real(c_double), parameter :: hsdc_tol = real(1.0e-17, c_double)

contains

   ! stable merge permutation of two sorted runs (LAPACK DLAMRG semantics: ties to
   ! the first run; strides +1 or -1 per run)
   subroutine hsdc_lamrg(n1, n2, a, s1, s2, index)
      integer, intent(in) :: n1, n2, s1, s2
      double precision, intent(in) :: a(*)
      integer, intent(out) :: index(*)
      integer :: i, ind1, ind2, m1, m2
      m1 = n1
      m2 = n2
      if (s1 > 0) then
         ind1 = 1
      else
         ind1 = n1
      end if
      if (s2 > 0) then
         ind2 = 1 + n1
      else
         ind2 = n1 + n2
      end if
      i = 1
      do while (m1 > 0 .and. m2 > 0)
         if (a(ind1) <= a(ind2)) then
            index(i) = ind1
            ind1 = ind1 + s1
            m1 = m1 - 1
         else
            index(i) = ind2
            ind2 = ind2 + s2
            m2 = m2 - 1
         end if
         i = i + 1
      end do
      do while (m1 > 0)
         index(i) = ind1
         ind1 = ind1 + s1
         m1 = m1 - 1
         i = i + 1
      end do
      do while (m2 > 0)
         index(i) = ind2
         ind2 = ind2 + s2
         m2 = m2 - 1
         i = i + 1
      end do
   end subroutine

   ! library status -> INFO: runtime failures (HSDC_ERR_*, <= -1000) become the
   ! positive range 1000.. documented in include/hsdc.h
   integer function hsdc_info(rc)
      integer(c_int), intent(in) :: rc
      if (rc <= -1000) then
         hsdc_info = -rc
      else
         hsdc_info = rc
      end if
   end function
end module hsdc_c

! MDSTEDC (mdstedc.f90:139)
subroutine mdstedc(compz, n, d, e, z, ldz, work, lwork, iwork, liwork, info)
   use hsdc_c
   implicit none
   character :: compz
   integer :: n, ldz, lwork, liwork, info
   double precision :: d(*), e(*), z(ldz, *), work(*)
   integer :: iwork(*)
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   info = 0
   ctx = hsdc_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   rc = hsdc_stedc(ctx, compz, n, d, e, z, ldz, work, lwork, iwork, liwork)
   info = hsdc_info(rc)
   if (info < 0) call xerbla('DSTEDC', -info)
end subroutine

! MDLAED0 (mdlaed0.f90:121).  QSTORE/LDQS/WORK/IWORK/WFLAG are accepted and unused
! except for the LDQS check; WFLAG is set to 1 (the HSS branch is available).
subroutine mdlaed0(icompq, qsiz, n, d, e, q, ldq, qstore, ldqs, work, iwork, info, wflag)
   use hsdc_c
   implicit none
   integer :: icompq, qsiz, n, ldq, ldqs, info, wflag
   double precision :: d(*), e(*), q(ldq, *), qstore(ldqs, *), work(*)
   integer :: iwork(*)
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   info = 0
   if (icompq < 0 .or. icompq > 2) then
      info = -1
   else if (icompq == 1 .and. qsiz < max(0, n)) then
      info = -2
   else if (n < 0) then
      info = -3
   else if (ldq < max(1, n)) then
      info = -7
   else if (ldqs < max(1, n)) then
      info = -9
   end if
   if (info /= 0) then
      call xerbla('DLAED0', -info)
      return
   end if
   if (n == 0) return
   wflag = 1
   ctx = hsdc_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   rc = hsdc_laed0(ctx, icompq, qsiz, n, d, e, q, ldq)
   info = hsdc_info(rc)
end subroutine

! MDLAED2 (mdlaed2.f90:149): deflation on the device (hsdc_deflate), the Givens
! rotations applied to Q here in the order the scan made them, then the column
! packing of Q2 and Q of the dense (K < BSMLZ) or HSS branch.
subroutine mdlaed2(k, n, n1, d, q, ldq, indxq, rho, z, dlamda, w, q2, indx, indxc, indxp, coltyp, hflag, info)
   use hsdc_c
   implicit none
   integer :: k, n, n1, ldq, info
   double precision :: rho
   logical :: hflag
   integer :: indxq(*), indx(*), indxc(*), indxp(*), coltyp(*)
   double precision :: d(*), q(ldq, *), z(*), dlamda(*), w(*), q2(*)
   type(c_ptr) :: ctx
   integer(c_int) :: rc, kk, nrot
   real(c_double) :: rho_o
   integer, allocatable :: rpj(:), rnj(:)
   double precision, allocatable :: rc_(:), rs_(:), drot(:), dfl(:), col(:)
   integer :: i, j, js, n2, iq1, iq2, ct, r, ctot(4)
   info = 0
   hflag = .true.
   if (n < 0) then
      info = -2
   else if (ldq < max(1, n)) then
      info = -6
   else if (min(1, n / 2) > n1 .or. (n / 2) < n1) then
      info = -3
   end if
   if (info /= 0) then
      call xerbla('DLAED2', -info)
      return
   end if
   n2 = n - n1
   allocate(rpj(max(n, 1)), rnj(max(n, 1)), rc_(max(n, 1)), rs_(max(n, 1)), drot(max(n, 1)), dfl(max(n, 1)))
   dfl(1:n) = d(1:n)
   ctx = hsdc_thread_ctx()
   rc = -1003
   if (c_associated(ctx)) &
      rc = hsdc_deflate(ctx, n, n1, dfl, z, indxq, rho, hsdc_bsmlz, kk, rho_o, dlamda, w, indxp, indx, indxc, &
                        coltyp, nrot, rpj, rnj, rc_, rs_)
   if (rc /= 0) then
      info = hsdc_info(rc)
      return
   end if
   k = kk
   rho = rho_o
   if (k == 0) then
      ! everything deflated (mdlaed2.f90:273-285): D and the columns of Q in sorted order
      d(1:n) = dfl(1:n)
      dlamda(1:n) = d(1:n)
      do j = 1, n
         q2(1 + (j - 1) * n:j * n) = q(1:n, indx(j))
      end do
      do j = 1, n
         q(1:n, j) = q2(1 + (j - 1) * n:j * n)
      end do
      deallocate(rpj, rnj, rc_, rs_, drot, dfl)
      return
   end if
   ! rotations as DROT(N, Q(1,PJ), 1, Q(1,NJ), 1, C, S) (mdlaed2.f90:352)
   allocate(col(n))
   do r = 1, nrot
      do i = 1, n
         col(i) = rc_(r) * q(i, rpj(r)) + rs_(r) * q(i, rnj(r))
         q(i, rnj(r)) = rc_(r) * q(i, rnj(r)) - rs_(r) * q(i, rpj(r))
         q(i, rpj(r)) = col(i)
      end do
   end do
   ! value of every column after the scan: non-deflated in INDXP(1:K) order, deflated after
   do j = 1, k
      drot(indxp(j)) = dlamda(j)
   end do
   do j = k + 1, n
      drot(indxp(j)) = dfl(j)
   end do
   ctot(1:4) = coltyp(1:4)
   if (k < hsdc_bsmlz) then
      iq1 = 1
      iq2 = 1 + (ctot(1) + ctot(2)) * n1
      i = 1
      do ct = 1, 4
         do j = 1, ctot(ct)
            js = indx(i)
            if (ct == 1 .or. ct == 2) then
               q2(iq1:iq1 + n1 - 1) = q(1:n1, js)
               iq1 = iq1 + n1
            end if
            if (ct == 2 .or. ct == 3) then
               q2(iq2:iq2 + n2 - 1) = q(n1 + 1:n, js)
               iq2 = iq2 + n2
            end if
            if (ct == 4) then
               if (j == 1) iq1 = iq2
               q2(iq2:iq2 + n - 1) = q(1:n, js)
               iq2 = iq2 + n
            end if
            z(i) = drot(js)
            i = i + 1
         end do
      end do
      if (ctot(4) == 0) iq1 = iq2
      if (k < n) then
         do j = 1, n - k
            q(1:n, k + j) = q2(iq1 + (j - 1) * n:iq1 + j * n - 1)
         end do
         d(k + 1:n) = z(k + 1:n)
      end if
   else
      hflag = .false.
      do j = 1, n
         q2(1 + (j - 1) * n:j * n) = q(1:n, indx(j))
         z(j) = drot(indx(j))
      end do
      do j = 1, n
         q(1:n, j) = q2(1 + (j - 1) * n:j * n)
      end do
      d(k + 1:n) = z(k + 1:n)
   end if
   coltyp(1:4) = ctot(1:4)
   deallocate(rpj, rnj, rc_, rs_, drot, dfl, col)
end subroutine

! MDLAED3 (mdlaed3.f90:124): roots, Loewner z, DIFL/DIFR/ALPHA on the device, then the
! HSS eigenvector update Q(:,1:K) = Q2 * C (DHSSEVC with Ni = 200, TOL = 1e-17).
! K = 1, 2 (which the reference's chain never sends here) are solved in place.
subroutine mdlaed3(k, n, d, q, ldq, rho, dlamda, q2, z, difl, difr, alpha, work, info)
   use hsdc_c
   implicit none
   integer :: k, n, ldq, info
   double precision :: rho
   double precision :: d(*), q(ldq, *), dlamda(*), q2(n, *), z(*), difl(*), difr(*), alpha(*), work(*)
   type(c_ptr) :: ctx
   integer(c_int) :: rc, org(2), nit(2)
   double precision :: tau(2), lam(2), v(2, 2), s
   integer :: i, j
   info = 0
   if (k < 0) then
      info = -1
   else if (n < k) then
      info = -2
   else if (ldq < max(1, n)) then
      info = -6
   end if
   if (info /= 0) then
      call xerbla('DLAED3', -info)
      return
   end if
   if (k == 0) return
   ctx = hsdc_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   if (k >= 3) then
      rc = hsdc_loewner(ctx, k, rho, dlamda, z, d, difl, difr, alpha)
      if (rc == 0) rc = hsdc_hss_evc(ctx, k, hsdc_ni, hsdc_tol, dlamda, d, z, alpha, difl, difr, q2, n, n, &
                                     c_null_ptr)
      if (rc == 0) then
         do j = 1, k
            q(1:n, j) = q2(1:n, j)
         end do
      end if
   else
      rc = hsdc_secular(ctx, k, dlamda, z, rho, lam, org, tau, nit)
      if (rc == 0) then
         ! eigenvectors of diag(dlamda) + rho z z': v_j(i) = z(i) / (dlamda(i) - lam(j))
         do j = 1, k
            s = 0
            do i = 1, k
               v(i, j) = z(i) / ((dlamda(i) - dlamda(org(j))) - tau(j))
               s = s + v(i, j)**2
            end do
            v(1:k, j) = v(1:k, j) / sqrt(s)
            d(j) = lam(j)
         end do
         do j = 1, k
            do i = 1, n
               work(i) = sum(q2(i, 1:k) * v(1:k, j))
            end do
            q(1:n, j) = work(1:n)
         end do
      end if
   end if
   info = hsdc_info(rc)
end subroutine

! MDLAED1 (mdlaed1.f90:109): one merge.  z from the last row of Q1 and the first row
! of Q2, MDLAED2, then DLAED3 (dense branch, on the device) or MDLAED3 (HSS branch),
! then INDXQ = DLAMRG(K, N-K, D, 1, -1).  WORK/IWORK hold the same intermediate
! arrays as the reference (WORK >= 4N + N^2 + 3N for the HSS branch).
subroutine mdlaed1(n, d, q, ldq, indxq, rho, cutpnt, work, iwork, info, wflag)
   use hsdc_c
   implicit none
   integer :: n, ldq, cutpnt, info, wflag
   double precision :: rho
   integer :: indxq(*), iwork(*)
   double precision :: d(*), q(ldq, *), work(*)
   integer :: iz, idlmda, iw, iq2, indx, indxc, coltyp, indxp, k, i, is, idifl, idifr, ialpha
   logical :: hflag
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   info = 0
   if (n < 0) then
      info = -1
   else if (ldq < max(1, n)) then
      info = -4
   else if (min(1, n / 2) > cutpnt .or. (n / 2) < cutpnt) then
      info = -7
   end if
   if (info /= 0) then
      call xerbla('DLAED1', -info)
      return
   end if
   if (n == 0) return
   iz = 1
   idlmda = iz + n
   iw = idlmda + n
   iq2 = iw + n
   indx = 1
   indxc = indx + n
   coltyp = indxc + n
   indxp = coltyp + n
   do i = 1, cutpnt
      work(iz + i - 1) = q(cutpnt, i)
   end do
   do i = cutpnt + 1, n
      work(iz + i - 1) = q(cutpnt + 1, i)
   end do
   call mdlaed2(k, n, cutpnt, d, q, ldq, indxq, rho, work(iz), work(idlmda), work(iw), work(iq2), &
                iwork(indx), iwork(indxc), iwork(indxp), iwork(coltyp), hflag, info)
   if (info /= 0) return
   if (k /= 0) then
      if (hflag) then
         ctx = hsdc_thread_ctx()
         rc = -1003
         if (c_associated(ctx)) &
            rc = hsdc_laed3(ctx, k, n, cutpnt, d, q, ldq, rho, work(idlmda), work(iq2), iwork(indxc), &
                            iwork(coltyp), work(iw))
         info = hsdc_info(rc)
      else
         wflag = 1
         idifl = k * n + iq2
         idifr = idifl + k
         ialpha = idifr + k
         is = ialpha + k
         call mdlaed3(k, n, d, q, ldq, rho, work(idlmda), work(iq2), work(iw), work(idifl), work(idifr), &
                      work(ialpha), work(is), info)
      end if
      if (info /= 0) return
      call hsdc_lamrg(k, n - k, d, 1, -1, indxq)
   else
      do i = 1, n
         indxq(i) = i
      end do
   end if
end subroutine


! MDBDSDC (mdbdsdc.f90:1-2).  COMPQ 'N' or 'I'; 'P' (the compact form) returns
! INFO = -2.  Q/IQ/WORK/IWORK are accepted and unused.
subroutine mdbdsdc(uplo, compq, n, d, e, u, ldu, vt, ldvt, q, iq, work, iwork, info)
   use hsdc_c
   implicit none
   character :: uplo, compq
   integer :: n, ldu, ldvt, info
   double precision :: d(*), e(*), u(ldu, *), vt(ldvt, *), q(*), work(*)
   integer :: iq(*), iwork(*)
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   info = 0
   ctx = hsdc_bd_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   rc = hsdc_bdsdc(ctx, uplo, compq, n, d, e, u, ldu, vt, ldvt, q, iq, work, iwork)
   info = hsdc_info(rc)
   if (info < 0) call xerbla('DBDSDC', -info)
end subroutine


! CauchyHssMM (cauchyhssmm.f90:1-123): B(1:N, 1:N) := A * B for the Cauchy-like
! A(i,j) = U(i) V(j) / (D(i) - F(j)), A put in HSS form by cauchy2hss with leaf size
! NI and tolerance TOL, NCOL columns of B multiplied at a time (or more); TIME(1) and
! TIME(2) receive the construction and multiplication times in seconds.  D, F, U, V
! are left unchanged (the reference destroys them).  The reference has no INFO: an
! illegal argument, or D and F sharing a value, stops with a message, as its failed
! allocations do.
subroutine cauchyhssmm(d, f, u, v, n, ni, tol, ncol, b, ldb, time)
   use hsdc_c
   implicit none
   integer :: n, ni, ncol, ldb
   double precision :: tol
   double precision :: d(*), f(*), u(*), v(*), b(ldb, *), time(2)
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   ctx = hsdc_thread_ctx()
   if (.not. c_associated(ctx)) then
      write (*, *) 'cauchyhssmm: no device context'
      stop
   end if
   rc = hsdc_cauchyhssmm(ctx, d, f, u, v, n, ni, tol, ncol, b, ldb, time, c_null_ptr)
   if (rc /= 0) then
      write (*, *) 'cauchyhssmm failed: ', rc
      stop
   end if
end subroutine


! MDSBEVD (mdsbevd.f90:183-184): all eigenvalues (ascending, in D) and, for COMPZ 'I',
! the eigenvectors (Q) of the symmetric band matrix in AB (LDAB >= KD + 1, UPLO 'U' or
! 'L' storage).  WORK/IWORK are accepted and unused.  INFO as the C entry hsdc_sbevd:
! -i for an illegal argument i (XERBLA is called with the reference's name), 1 when a
! leaf, a split or a secular root did not converge.
subroutine mdsbevd(compz, uplo, n, kd, ab, ldab, d, q, ldq, work, lwork, iwork, liwork, info)
   use hsdc_c
   implicit none
   character :: compz, uplo
   integer :: n, kd, ldab, ldq, lwork, liwork, info
   double precision :: ab(ldab, *), d(*), q(ldq, *), work(*)
   integer :: iwork(*)
   type(c_ptr) :: ctx
   integer(c_int) :: rc
   info = 0
   ctx = hsdc_sb_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   rc = hsdc_sbevd(ctx, compz, uplo, n, kd, ab, ldab, d, q, ldq, work, lwork, iwork, liwork)
   info = hsdc_info(rc)
   if (info < 0) call xerbla('MDSBEVD', -info)
end subroutine


! DHSSMM (DHSSMM.f90:1-125): B(1:N, 1:K) := A * B with the dense A (LDA x N, M = N) put
! in HSS form by truncated SVDs (leaf size NI, tolerance TOL), NCOL columns of B at a time
! or more; TIME(1:2) construction and multiplication seconds.  PRANK is unused (as in the
! reference's mat2hssvd); A is left unchanged.  INFO: 0, -i for an illegal argument i, or
! 1000 + code for a runtime failure (the reference never sets it).
subroutine dhssmm(a, lda, m, n, b, ldb, k, tol, prank, ni, ncol, time, info)
   use hsdc_c
   implicit none
   integer :: lda, m, n, ldb, k, prank, ni, ncol, info
   double precision :: tol
   double precision :: a(lda, *), b(ldb, *), time(2)
   type(c_ptr) :: ctx
   ctx = hsdc_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   info = hsdc_info(hsdc_dhssmm(ctx, a, lda, m, n, b, ldb, k, tol, prank, ni, ncol, time, c_null_ptr))
end subroutine

! DHSSMML (DHSSMML.f90:2-122): B(1:K, 1:N) := B * A, otherwise as DHSSMM.
subroutine dhssmml(a, lda, m, n, b, ldb, k, tol, prank, ni, ncol, time, info)
   use hsdc_c
   implicit none
   integer :: lda, m, n, ldb, k, prank, ni, ncol, info
   double precision :: tol
   double precision :: a(lda, *), b(ldb, *), time(2)
   type(c_ptr) :: ctx
   ctx = hsdc_thread_ctx()
   if (.not. c_associated(ctx)) then
      info = 1003
      return
   end if
   info = hsdc_info(hsdc_dhssmml(ctx, a, lda, m, n, b, ldb, k, tol, prank, ni, ncol, time, c_null_ptr))
end subroutine
