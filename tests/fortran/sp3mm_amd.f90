! ISO_C_BINDING interface of libsp3mm_amd's 1-based (Fortran) entry points.
! The derived types have the byte layout of the reference's structs:
! spmat is code/src/include/sparseMatrix.h:41-53 (idx_t = unsigned long,
! read as c_long), sp3mm_config is code/src/include/config.h:42-49.
module sp3mm_amd
  use iso_c_binding
  implicit none

  type, bind(c) :: spmat
     integer(c_long) :: nz, m, n
     type(c_ptr)     :: as, ja, irp
     integer(c_long) :: max_row_nz
  end type spmat

  type, bind(c) :: sp3mm_config
     integer(c_short) :: grid_rows, grid_cols
     integer(c_int)   :: symb_mm_row_impl_id   ! RBTREE = 0, IDXMAP = 1
     integer(c_int)   :: thread_num
     type(c_ptr)      :: chunk_distrb_func
  end type sp3mm_config

  interface
     ! spmat* sp3mmRowByRowPair_1(spmat*, spmat*, spmat*, CONFIG*, SPMM_INTERF)
     type(c_ptr) function sp3mm_rap_pair_1(r, ac, p, cfg, spmm) bind(c, name="sp3mmRowByRowPair_1")
       import :: c_ptr, c_funptr, spmat, sp3mm_config
       type(spmat), intent(in) :: r, ac, p
       type(sp3mm_config), intent(in) :: cfg
       type(c_funptr), value :: spmm      ! c_funloc of an spmm*_1, or c_null_funptr
     end function sp3mm_rap_pair_1

     ! spmat* sp3mmRowByRowMerged_1(spmat*, spmat*, spmat*, CONFIG*, SPMM_INTERF)
     type(c_ptr) function sp3mm_rap_merged_1(r, ac, p, cfg, spmm) bind(c, name="sp3mmRowByRowMerged_1")
       import :: c_ptr, c_funptr, spmat, sp3mm_config
       type(spmat), intent(in) :: r, ac, p
       type(sp3mm_config), intent(in) :: cfg
       type(c_funptr), value :: spmm      ! ignored, as in the reference
     end function sp3mm_rap_merged_1

     ! spmat* spmmRowByRow_1(spmat*, spmat*, CONFIG*) and the symbolic-numeric form
     type(c_ptr) function spmm_row_by_row_1(a, b, cfg) bind(c, name="spmmRowByRow_1")
       import :: c_ptr, spmat, sp3mm_config
       type(spmat), intent(in) :: a, b
       type(sp3mm_config), intent(in) :: cfg
     end function spmm_row_by_row_1
     type(c_ptr) function spmm_row_by_row_symbnum_1(a, b, cfg) bind(c, name="spmmRowByRow_SymbNum_1")
       import :: c_ptr, spmat, sp3mm_config
       type(spmat), intent(in) :: a, b
       type(sp3mm_config), intent(in) :: cfg
     end function spmm_row_by_row_symbnum_1

     ! void freeSpmat(spmat*)
     subroutine sp3mm_free_spmat(a) bind(c, name="freeSpmat")
       import :: c_ptr
       type(c_ptr), value :: a
     end subroutine sp3mm_free_spmat
  end interface
end module sp3mm_amd
