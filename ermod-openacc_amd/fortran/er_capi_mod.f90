! er_capi_mod.f90 -- iso_c_binding interface of include/er_capi.h and
! include/er_traj.h for a Fortran erdst host.  engconst/engstore bind the GPU hot path through this
! module; every argument is an engmain array or scalar passed as is
! (1-based indices, column-major arrays).
module er_capi
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ER_OK = 0, ER_E_ARG = 1, ER_E_UNSUPPORTED = 2, &
       ER_E_HIP = 3, ER_E_HALT = 4, ER_E_STATE = 5
  integer(c_int), parameter, public :: ER_PLAIN = 0, ER_BATCHED = 1, ER_PLAIN_ASYNC = 2
  integer(c_int), parameter, public :: ER_TRAJ_EOF = 100, ER_TRAJ_IO = 101, ER_TRAJ_FORMAT = 102

  type, bind(C), public :: er_topology
     integer(c_int32_t) :: numatm, nummol, ljtype_max, numslv
     type(c_ptr) :: mol_begin_index, ljtype, uvspec, charge, ljene_mat, ljlensq_mat
  end type er_topology

  type, bind(C), public :: er_params
     real(c_float) :: elecut, lwljcut, upljcut, screen
     integer(c_int32_t) :: ljswitch, cltype, splodr
     integer(c_int32_t) :: ms(3)
     integer(c_int32_t) :: boxshp, slttype, wgtslf, estype, corrcal, selfcal
     real(c_double) :: temp
     integer(c_int32_t) :: force_calculation
  end type er_params

  public :: er_create, er_set_mesh, er_set_solute, er_set_solutes, er_select_solute, er_set_solute_molecules, &
       er_set_frame, er_set_insertions, er_run, er_sync, er_get_uvengy, er_get_nins, er_fetch, &
       er_get_flceng, er_clear, er_last_error, er_destroy, er_traj_open, er_traj_read, er_traj_close, &
       er_autotune, er_set_tuning, er_get_tuning, er_set_cell_width, er_get_cell_width

  interface
     integer(c_int) function er_create(topo, prm, ctx) bind(C, name="er_create")
       import :: c_int, c_ptr, er_topology, er_params
       type(er_topology), intent(in) :: topo
       type(er_params), intent(in) :: prm
       type(c_ptr), intent(out) :: ctx
     end function er_create

     integer(c_int) function er_set_mesh(ctx, uvmax, uvsoft, uvcrd, esmax, escrd) &
          bind(C, name="er_set_mesh")
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: uvmax(*), uvsoft(*)
       real(c_double), intent(in) :: uvcrd(*)
       integer(c_int32_t), value :: esmax
       real(c_double), intent(in) :: escrd(*)
     end function er_set_mesh

     integer(c_int) function er_set_solute(ctx, nsite, ljtype, charge) bind(C, name="er_set_solute")
       import :: c_int, c_int32_t, c_float, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsite
       integer(c_int32_t), intent(in) :: ljtype(*)
       real(c_float), intent(in) :: charge(*)
     end function er_set_solute

     integer(c_int) function er_set_solutes(ctx, nsolute, nsite, ljtype, charge) bind(C, name="er_set_solutes")
       import :: c_int, c_int32_t, c_float, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsolute
       integer(c_int32_t), intent(in) :: nsite(*), ljtype(*)
       real(c_float), intent(in) :: charge(*)
     end function er_set_solutes

     integer(c_int) function er_select_solute(ctx, k) bind(C, name="er_select_solute")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
     end function er_select_solute

     integer(c_int) function er_set_solute_molecules(ctx, nslt, tagslt) &
          bind(C, name="er_set_solute_molecules")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nslt
       integer(c_int32_t), intent(in) :: tagslt(*)
     end function er_set_solute_molecules

     integer(c_int) function er_set_frame(ctx, sitepos, cell, stat_weight_system) &
          bind(C, name="er_set_frame")
       import :: c_int, c_float, c_ptr
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: sitepos(3, *), cell(3, 3)
       real(c_float), value :: stat_weight_system
     end function er_set_frame

     integer(c_int) function er_set_insertions(ctx, nins, sltpos, stat_weight_solute) &
          bind(C, name="er_set_insertions")
       import :: c_int, c_int32_t, c_float, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nins
       real(c_float), intent(in) :: sltpos(3, *)
       real(c_float), intent(in) :: stat_weight_solute(*)
     end function er_set_insertions

     integer(c_int) function er_run(ctx, mode) bind(C, name="er_run")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mode
     end function er_run

     integer(c_int) function er_sync(ctx) bind(C, name="er_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function er_sync

     integer(c_int) function er_get_uvengy(ctx, uvengy, uvengy0) bind(C, name="er_get_uvengy")
       import :: c_int, c_float, c_ptr
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: uvengy(*), uvengy0(*)
     end function er_get_uvengy

     integer(c_int) function er_get_nins(ctx, nins) bind(C, name="er_get_nins")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: nins
     end function er_get_nins

     integer(c_int) function er_fetch(ctx, edens, ecorr, eself, scal, minuv, maxuv, slnuv) &
          bind(C, name="er_fetch")
       import :: c_int, c_double, c_float, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: edens(*), ecorr(*), eself(*), scal(4), slnuv(*)
       real(c_float), intent(out) :: minuv(*), maxuv(*)
     end function er_fetch

     integer(c_int) function er_get_flceng(ctx, flceng) bind(C, name="er_get_flceng")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: flceng(*)
     end function er_get_flceng

     integer(c_int) function er_autotune(ctx, reps) bind(C, name="er_autotune")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: reps
     end function er_autotune

     integer(c_int) function er_set_tuning(ctx, chunk, pair_split, hist_slices) bind(C, name="er_set_tuning")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: chunk, pair_split, hist_slices
     end function er_set_tuning

     integer(c_int) function er_get_tuning(ctx, chunk, pair_split, hist_slices) bind(C, name="er_get_tuning")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: chunk, pair_split, hist_slices
     end function er_get_tuning

     integer(c_int) function er_set_cell_width(ctx, angstrom) bind(C, name="er_set_cell_width")
       import :: c_int, c_float, c_ptr
       type(c_ptr), value :: ctx
       real(c_float), value :: angstrom
     end function er_set_cell_width

     integer(c_int) function er_get_cell_width(ctx, angstrom) bind(C, name="er_get_cell_width")
       import :: c_int, c_float, c_ptr
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: angstrom
     end function er_get_cell_width

     integer(c_int) function er_clear(ctx) bind(C, name="er_clear")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function er_clear

     integer(c_int) function er_last_error(ctx, code, species, energy) bind(C, name="er_last_error")
       import :: c_int, c_int32_t, c_float, c_char, c_ptr
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(out) :: code(8)
       integer(c_int32_t), intent(out) :: species
       real(c_float), intent(out) :: energy
     end function er_last_error

     subroutine er_destroy(ctx) bind(C, name="er_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine er_destroy

     ! open_trajectory / read_trajectory / close_trajectory (trajectory.F90:41-96)
     ! for DCD files; path is NUL-terminated (trim(name) // c_null_char)
     integer(c_int) function er_traj_open(path, traj, natoms) bind(C, name="er_traj_open")
       import :: c_int, c_int32_t, c_char, c_ptr
       character(kind=c_char), intent(in) :: path(*)
       type(c_ptr), intent(out) :: traj
       integer(c_int32_t), intent(out) :: natoms
     end function er_traj_open

     integer(c_int) function er_traj_read(traj, natoms, xyz, cell) bind(C, name="er_traj_read")
       import :: c_int, c_int32_t, c_float, c_ptr
       type(c_ptr), value :: traj
       integer(c_int32_t), value :: natoms
       real(c_float), intent(out) :: xyz(*), cell(*)
     end function er_traj_read

     subroutine er_traj_close(traj) bind(C, name="er_traj_close")
       import :: c_ptr
       type(c_ptr), value :: traj
     end subroutine er_traj_close
  end interface
end module er_capi
