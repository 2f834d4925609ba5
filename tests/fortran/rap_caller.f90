! A Fortran caller of the 1-based entry points, as an AMG4PSBLAS-style code
! would call them: the 2D 5-point Laplacian on an n x n grid (1-based CSR,
! sorted columns), piecewise-constant aggregation of 2 x 2 blocks (P), R = P^T.
! The Galerkin operator R*A*P then has the closed form C(I,J) = sum of A(i,j)
! over i in aggregate I and j in aggregate J (small integers, exact in FP64),
! and its pattern is every aggregate pair joined by an entry of A.  The pair
! form (with the library's SpMMs and with none) and the merged form are
! checked against it; the program prints "fortran caller ok" and stops with
! code 0 only if every product matches.
program rap_caller
  use iso_c_binding
  use sp3mm_amd
  implicit none
  integer, parameter :: n = 16, nc = n / 2, nf = n * n, ncoarse = nc * nc
  integer(c_long), target :: a_irp(nf + 1), a_ja(5 * nf), p_irp(nf + 1), p_ja(nf)
  integer(c_long), target :: r_irp(ncoarse + 1), r_ja(nf)
  real(c_double), target :: a_as(5 * nf), p_as(nf), r_as(nf)
  real(c_double) :: cref(ncoarse, ncoarse)
  logical :: cstruct(ncoarse, ncoarse)
  integer :: agg(nf), x, y, i, j, k, e, fails, fill(ncoarse)
  type(spmat) :: rm, am, pm
  type(sp3mm_config) :: cfg

  ! A: rows in grid order, neighbours (x,y-1), (x-1,y), self, (x+1,y), (x,y+1)
  e = 0
  a_irp(1) = 1
  do y = 0, n - 1
     do x = 0, n - 1
        i = y * n + x + 1
        if (y > 0)     call put(i - n, -1.0d0)
        if (x > 0)     call put(i - 1, -1.0d0)
        call put(i, 4.0d0)
        if (x < n - 1) call put(i + 1, -1.0d0)
        if (y < n - 1) call put(i + n, -1.0d0)
        a_irp(i + 1) = e + 1
        agg(i) = (y / 2) * nc + x / 2 + 1
     end do
  end do
  ! P: one 1.0 per fine row, in its aggregate's column; R = P^T
  do i = 1, nf
     p_irp(i) = i
     p_ja(i) = agg(i)
     p_as(i) = 1.0d0
  end do
  p_irp(nf + 1) = nf + 1
  r_irp(1) = 1
  do j = 1, ncoarse
     r_irp(j + 1) = r_irp(j) + count(agg == j)
  end do
  fill = 0
  do i = 1, nf
     k = int(r_irp(agg(i))) + fill(agg(i))
     r_ja(k) = i
     r_as(k) = 1.0d0
     fill(agg(i)) = fill(agg(i)) + 1
  end do
  ! closed form of R*A*P
  cref = 0.0d0
  cstruct = .false.
  do i = 1, nf
     do k = int(a_irp(i)), int(a_irp(i + 1)) - 1
        cref(agg(i), agg(a_ja(k))) = cref(agg(i), agg(a_ja(k))) + a_as(k)
        cstruct(agg(i), agg(a_ja(k))) = .true.
     end do
  end do

  rm = spmat(int(nf, c_long), int(ncoarse, c_long), int(nf, c_long), c_loc(r_as), c_loc(r_ja), c_loc(r_irp), &
             0_c_long)
  am = spmat(int(e, c_long), int(nf, c_long), int(nf, c_long), c_loc(a_as), c_loc(a_ja), c_loc(a_irp), 0_c_long)
  pm = spmat(int(nf, c_long), int(nf, c_long), int(ncoarse, c_long), c_loc(p_as), c_loc(p_ja), c_loc(p_irp), 0_c_long)
  cfg = sp3mm_config(20_c_short, 2_c_short, 1_c_int, 4_c_int, c_null_ptr)

  fails = 0
  fails = fails + check("sp3mmRowByRowPair_1 + spmmRowByRow_1", &
                        sp3mm_rap_pair_1(rm, am, pm, cfg, c_funloc(spmm_row_by_row_1)))
  fails = fails + check("sp3mmRowByRowPair_1 + spmmRowByRow_SymbNum_1", &
                        sp3mm_rap_pair_1(rm, am, pm, cfg, c_funloc(spmm_row_by_row_symbnum_1)))
  fails = fails + check("sp3mmRowByRowPair_1 (no SpMM given)", sp3mm_rap_pair_1(rm, am, pm, cfg, c_null_funptr))
  fails = fails + check("sp3mmRowByRowMerged_1", sp3mm_rap_merged_1(rm, am, pm, cfg, c_null_funptr))
  ! the inputs are borrowed, not modified (still 1-based)
  if (a_irp(1) /= 1 .or. r_irp(1) /= 1 .or. p_ja(1) /= agg(1)) then
     print *, "inputs were modified"
     fails = fails + 1
  end if
  if (fails /= 0) then
     print *, "fortran caller FAILED:", fails
     stop 1
  end if
  print '(a)', "fortran caller ok"

contains

  subroutine put(col, v)
    integer, intent(in) :: col
    real(c_double), intent(in) :: v
    e = e + 1
    a_ja(e) = col
    a_as(e) = v
  end subroutine put

  ! 0 if the returned 1-based CSR equals the closed form, else 1; frees it
  integer function check(what, res) result(bad)
    character(*), intent(in) :: what
    type(c_ptr), intent(in) :: res
    type(spmat), pointer :: c
    integer(c_long), pointer :: irp(:), ja(:)
    real(c_double), pointer :: as(:)
    integer :: row, q, nrow
    bad = 1
    if (.not. c_associated(res)) then
       print *, what, ": NULL result"
       return
    end if
    call c_f_pointer(res, c)
    if (c%m /= ncoarse .or. c%n /= ncoarse .or. c%nz /= count(cstruct)) then
       print *, what, ": shape or nnz", c%m, c%n, c%nz
       call sp3mm_free_spmat(res)
       return
    end if
    call c_f_pointer(c%irp, irp, [ncoarse + 1])
    call c_f_pointer(c%ja, ja, [int(c%nz)])
    call c_f_pointer(c%as, as, [int(c%nz)])
    if (irp(1) /= 1 .or. irp(ncoarse + 1) /= c%nz + 1) then
       print *, what, ": row offsets are not 1-based"
       call sp3mm_free_spmat(res)
       return
    end if
    do row = 1, ncoarse
       nrow = int(irp(row + 1) - irp(row))
       if (nrow /= count(cstruct(row, :))) then
          print *, what, ": row", row, "has", nrow, "entries"
          call sp3mm_free_spmat(res)
          return
       end if
       do q = int(irp(row)), int(irp(row + 1)) - 1
          if (q > int(irp(row))) then
             if (ja(q) <= ja(q - 1)) then
                print *, what, ": row", row, "columns not increasing"
                call sp3mm_free_spmat(res)
                return
             end if
          end if
          if (.not. cstruct(row, ja(q)) .or. as(q) /= cref(row, ja(q))) then
             print *, what, ": entry", row, ja(q), as(q)
             call sp3mm_free_spmat(res)
             return
          end if
       end do
    end do
    call sp3mm_free_spmat(res)
    print '(a, a)', "ok: ", what
    bad = 0
  end function check
end program rap_caller
