! erdst_host.f90 -- a Fortran erdst driver over the C ABI: the engdiv x frame
! loop of trjmain (enganal.F90:86-95), engconst's device work through
! er_run, and engstore's normalisation and writers (engproc.F90:520-725).
!
! Input: one stream file (argument 1) holding the engmain topology, the
! &ene_param values, the energy-coordinate mesh and, per frame, the frame's
! coordinates and the insertion coordinates (layout: read_workload below).
! Optional arguments: "plain" (one insertion per launch sequence);
! "history=FILE" takes each frame's coordinates and cell from a DCD file
! instead, as erdst reads HISTORY (setconf.F90:103-105: a failed read halts
! with set_trj); "sltconf=FILE" takes the refs insertions from a DCD file of
! one solute configuration per frame, read in turn and reopened at its end
! as erdst reads SltConf (setconf.F90:106-116); "autotune" times the launch
! shapes on the first frame (er_autotune) before its er_run, which changes
! no output.
! Output (current directory): engref.XX / corref.XX (engsln / corsln for
! soln runs), slfeng.XX (selfcal), weight_refs|weight_soln, aveuv.tt (soln),
! uvrange.tt, as engstore writes them, and per frame flcuv.tt (soln) or
! progress.tt (refs), as engconst writes them (engproc.F90:308-315, 485-509).
program erdst_host
  use iso_c_binding
  use er_capi
  implicit none
  character(len=512) :: path, history, sltconf
  type(c_ptr) :: ctx, htraj, hslt
  type(er_topology) :: topo
  type(er_params) :: prm
  integer(c_int32_t) :: numatm, nummol, ntype, numslv, nslt, nins, nframes, ermax, engdiv, mode, esmax
  integer(c_int32_t), allocatable, target :: mol_begin(:), ljtype(:), uvspec(:), slt_lj(:), tagslt(:)
  integer(c_int32_t), allocatable :: uvmax(:), uvsoft(:)
  real(c_float), allocatable, target :: charge(:), ljene(:), ljlensq(:), slt_q(:)
  real(c_float), allocatable :: sitepos(:,:), sltpos(:,:), wslt(:), minuv(:), maxuv(:)
  real(c_float) :: cell(3,3)
  real(c_double), allocatable :: uvcrd(:), edens(:), ecorr(:), slnuv(:), avediv(:,:), escrd(:), eself(:)
  real(c_double), allocatable :: aveuv(:,:), flceng(:,:)
  real(c_double) :: scal(4), dummy(1)
  integer :: iu, rc, f, division, frames_per_div, k, io_flcuv, narg
  integer(c_int32_t) :: tune_chunk, tune_split, tune_slices
  logical :: autotune
  integer, parameter :: stdout = 6

  call get_command_argument(1, path)
  if (len_trim(path) == 0) stop "usage: erdst_host workload.bin [plain]"
  mode = ER_BATCHED
  history = ""; sltconf = ""; autotune = .false.
  do narg = 2, command_argument_count()
     block
       character(len=600) :: arg
       call get_command_argument(narg, arg)
       if (trim(arg) == "plain") then
          mode = ER_PLAIN
       else if (trim(arg) == "autotune") then
          autotune = .true.
       else if (arg(1:8) == "history=") then
          history = arg(9:)
       else if (arg(1:8) == "sltconf=") then
          sltconf = arg(9:)
       else
          stop "usage: erdst_host workload.bin [plain] [autotune] [history=FILE] [sltconf=FILE]"
       end if
     end block
  end do
  open(newunit=iu, file=trim(path), access="stream", form="unformatted", status="old", action="read")
  call read_workload()
  if (len_trim(history) > 0) call open_dcd(history, numatm, htraj)
  if (len_trim(sltconf) > 0) then
     if (prm%slttype == 1) stop "sltconf= applies to refs runs"
     call open_dcd(sltconf, nslt, hslt)
  end if

  topo%numatm = numatm; topo%nummol = nummol; topo%ljtype_max = ntype; topo%numslv = numslv
  topo%mol_begin_index = c_loc(mol_begin); topo%ljtype = c_loc(ljtype); topo%uvspec = c_loc(uvspec)
  topo%charge = c_loc(charge); topo%ljene_mat = c_loc(ljene); topo%ljlensq_mat = c_loc(ljlensq)
  rc = er_create(topo, prm, ctx)
  call check(rc, "er_create")
  if (esmax > 0) then
     rc = er_set_mesh(ctx, uvmax, uvsoft, uvcrd, esmax, escrd)
  else
     rc = er_set_mesh(ctx, uvmax, uvsoft, uvcrd, 0_c_int32_t, dummy)
  end if
  call check(rc, "er_set_mesh")
  if (prm%slttype == 1) then
     rc = er_set_solute_molecules(ctx, nins, tagslt)
  else
     rc = er_set_solute(ctx, nslt, slt_lj, slt_q)
  end if
  call check(rc, "er_set_solute")

  allocate(edens(ermax), ecorr(int(ermax, 8) * ermax), slnuv(numslv), minuv(0:numslv), maxuv(0:numslv))
  allocate(avediv(engdiv, 2), aveuv(engdiv, numslv), eself(max(esmax, 1)))
  ! engproc.F90:308-315: flcuv.tt (soln) or progress.tt (refs), status='new'
  if (prm%slttype == 1) then
     allocate(flceng(numslv, nins))
     open(newunit=io_flcuv, file='flcuv.tt', status='new', action='write')
  else
     open(newunit=io_flcuv, file='progress.tt', status='new', action='write')
  end if
  frames_per_div = nframes / engdiv
  do f = 1, nframes
     if (len_trim(history) > 0) then
        rc = er_traj_read(htraj, numatm, sitepos, cell)
        if (rc /= ER_OK) stop "set_trj: HISTORY has no further frame"
     else
        read(iu) sitepos
     end if
     rc = er_set_frame(ctx, sitepos, cell, 1.0_c_float)
     call check(rc, "er_set_frame")
     if (prm%slttype /= 1) then
        if (len_trim(sltconf) > 0) then
           call read_sltconf()
        else
           read(iu) sltpos
        end if
        rc = er_set_insertions(ctx, nins, sltpos, wslt)
        call check(rc, "er_set_insertions")
     end if
     if (autotune .and. f == 1) then
        rc = er_autotune(ctx, 0_c_int32_t)
        call check(rc, "er_autotune")
        rc = er_get_tuning(ctx, tune_chunk, tune_split, tune_slices)
        write(stdout, '(a,i0,a,i0,a,i0)') "erdst_host: launch shapes: chunk ", tune_chunk, ", pair split ", &
             tune_split, ", histogram slices ", tune_slices
     end if
     rc = er_run(ctx, mode)                                   ! engconst
     call check(rc, "er_run")
     rc = er_sync(ctx)                 ! engconst's halts and the mbin warning stop the run here
     call check(rc, "er_run")
     call frame_output(f)
     if (mod(f, frames_per_div) == 0) then
        division = f / frames_per_div
        call engstore(division)
        rc = er_clear(ctx)                                    ! engclear
        call check(rc, "er_clear")
     end if
  end do
  close(iu)
  close(io_flcuv)
  if (len_trim(history) > 0) call er_traj_close(htraj)
  if (len_trim(sltconf) > 0) call er_traj_close(hslt)
  call er_destroy(ctx)

contains

  subroutine open_dcd(name, natoms_expected, h)
    character(len=*), intent(in) :: name
    integer(c_int32_t), intent(in) :: natoms_expected
    type(c_ptr), intent(out) :: h
    integer(c_int32_t) :: n
    if (er_traj_open(trim(name) // c_null_char, h, n) /= ER_OK) &
         stop "unable to open trajectory"
    if (n /= natoms_expected) stop "# of atoms in trajectory does not match with # of atoms in configurations"
  end subroutine open_dcd

  ! one solute configuration per insertion, reopening SltConf at its end
  subroutine read_sltconf()
    integer :: n
    real(c_float) :: scell(9)
    do n = 0, nins - 1
       rc = er_traj_read(hslt, nslt, sltpos(:, n * nslt + 1:), scell)
       if (rc /= ER_OK) then
          call er_traj_close(hslt)
          call open_dcd(sltconf, nslt, hslt)
          rc = er_traj_read(hslt, nslt, sltpos(:, n * nslt + 1:), scell)
          if (rc /= ER_OK) stop "Failed to reload solute trajectory"
       end if
    end do
  end subroutine read_sltconf

  subroutine check(rc, what)
    integer, intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: code(8)
    integer(c_int32_t) :: sp
    real(c_float) :: en
    integer :: i, r2
    character(len=7) :: c7
    if (rc == ER_OK) return
    r2 = er_last_error(ctx, code, sp, en)
    do i = 1, 7
       c7(i:i) = code(i)
    end do
    write(stdout, '(A,A,A,I0)') " ", what, " failed, status ", rc
    if (rc == ER_E_HALT) then
       write(stdout, '(A,g12.4,A,i3,A)') '  energy of ', en, ' for ', sp, '-th species'
       write(stdout, '(A,A)') " halt code ", c7
    end if
    stop 1
  end subroutine check

  subroutine read_workload()
    integer(c_int32_t) :: hdr(22)
    real(c_float) :: rp(4)
    real(c_double) :: t
    read(iu) hdr
    numatm = hdr(1); nummol = hdr(2); ntype = hdr(3); numslv = hdr(4); nslt = hdr(5)
    nins = hdr(6); nframes = hdr(7); ermax = hdr(8)
    prm%slttype = hdr(9); prm%ms = hdr(10:12); prm%splodr = hdr(13); prm%ljswitch = hdr(14)
    prm%estype = hdr(15); prm%corrcal = hdr(16); prm%wgtslf = hdr(17); prm%selfcal = hdr(18)
    engdiv = hdr(19); esmax = hdr(20)
    if (esmax > 0 .and. prm%selfcal /= 1) stop "workload: a self-energy mesh needs selfcal = 1"
    if (esmax <= 0 .and. prm%selfcal == 1) stop "workload: selfcal = 1 needs a self-energy mesh"
    prm%cltype = hdr(21); prm%boxshp = hdr(22)
    read(iu) rp, t
    prm%elecut = rp(1); prm%lwljcut = rp(2); prm%upljcut = rp(3); prm%screen = rp(4); prm%temp = t
    prm%force_calculation = 0                   ! force_calculation = .false. (setconf.F90 default)
    allocate(mol_begin(nummol + 1), ljtype(numatm), uvspec(nummol), charge(numatm))
    allocate(ljene(ntype * ntype), ljlensq(ntype * ntype), uvmax(numslv), uvsoft(numslv), uvcrd(ermax))
    read(iu) mol_begin, ljtype, uvspec, charge, ljene, ljlensq, uvmax, uvsoft, uvcrd
    allocate(escrd(max(esmax, 1)))
    if (esmax > 0) read(iu) escrd(1:esmax)
    if (prm%slttype == 1) then
       allocate(tagslt(nins))
       read(iu) tagslt
    else
       allocate(slt_lj(nslt), slt_q(nslt))
       read(iu) slt_lj, slt_q
       allocate(sltpos(3, nslt * nins), wslt(nins))
       wslt(:) = 1.0
    end if
    read(iu) cell
    allocate(sitepos(3, numatm))
  end subroutine read_workload

  ! representative_bin_info (engproc.F90:1355-1387)
  subroutine bin_info(iduv, engleft, engmiddle, enginvwidth, spec)
    integer, intent(in) :: iduv
    real(c_double), intent(out) :: engleft, engmiddle, enginvwidth
    integer, intent(out) :: spec
    integer :: idmin, idpt, cnt
    idmin = 0
    do cnt = 1, numslv
       idpt = idmin + uvmax(cnt)
       if (iduv <= idpt) exit
       idmin = idpt
    end do
    spec = cnt
    idpt = iduv - idmin
    engleft = uvcrd(iduv)
    if (idpt <= uvsoft(spec) .and. iduv < ermax) then
       engmiddle = (uvcrd(iduv) + uvcrd(iduv + 1)) / 2.0
       enginvwidth = 1. / (uvcrd(iduv + 1) - uvcrd(iduv))
    else if (idpt < uvmax(spec)) then
       engmiddle = sqrt(uvcrd(iduv) * uvcrd(iduv + 1))
       enginvwidth = 1. / (uvcrd(iduv + 1) - uvcrd(iduv))
    else
       engmiddle = uvcrd(iduv)
       enginvwidth = 0.
    end if
  end subroutine bin_info

  ! engconst's per-frame output (engproc.F90:461-509, single process): per
  ! solute molecule of a soln frame its summed pair energy with each solvent
  ! species, written as real like the reference's flceng; the frame number for
  ! refs.  The workload's frames are the ones kept, so skpcnf = 1.
  subroutine frame_output(stnum)
    integer, intent(in) :: stnum
    integer :: cnt
    if (prm%slttype == 1) then
       rc = er_get_flceng(ctx, flceng)
       call check(rc, "er_get_flceng")
       do cnt = 1, nins
          if (nins == 1) then
             write(io_flcuv, '(i9, 9999f15.5)') stnum, real(flceng(1:numslv, cnt))
          else
             write(io_flcuv, '(2i9, 9999f15.5)') cnt, stnum, real(flceng(1:numslv, cnt))
          end if
       end do
    else
       write(io_flcuv, *) stnum
    end if
    flush(io_flcuv)
  end subroutine frame_output

  ! representative_bin_selfenergy (engproc.F90:1344-1353)
  real function selfenergy_coord(iduv)
    integer, intent(in) :: iduv
    if (iduv < esmax) then
       selfenergy_coord = (escrd(iduv) + escrd(iduv + 1)) / 2.0
    else
       selfenergy_coord = escrd(esmax)
    end if
  end function selfenergy_coord

  ! engstore for one block (engproc.F90:520-725, single process)
  subroutine engstore(division)
    integer, intent(in) :: division
    character(len=3) :: suffeng
    character(len=10), parameter :: numbers = '0123456789'
    real(c_double) :: engnorm, engsmpl, avslf, voffset, leftbin, middlebin, invwt
    integer :: iduv, pti, io, j, kk
    rc = er_fetch(ctx, edens, ecorr, eself, scal, minuv, maxuv, slnuv)
    call check(rc, "er_fetch")
    rc = er_sync(ctx)
    call check(rc, "er_sync")
    engnorm = scal(1); engsmpl = scal(2); avslf = scal(3); voffset = scal(4)
    edens(:) = edens(:) / engnorm
    if (prm%corrcal == 1) ecorr(:) = ecorr(:) / engnorm
    if (esmax > 0) eself(1:esmax) = eself(1:esmax) / engnorm
    avslf = avslf / engnorm
    avediv(division, 1) = engnorm / engsmpl
    if (prm%slttype == 1) then
       aveuv(division, :) = slnuv(:) / engnorm
       avediv(division, 2) = voffset - prm%temp * log(avslf)
    else
       avediv(division, 2) = voffset + prm%temp * log(avslf)
    end if
    if (division == engdiv) then
       if (prm%slttype == 1) then
          open(newunit=io, file='aveuv.tt', action='write')
          do kk = 1, engdiv
             write(io, '(i5,9999f15.5)') kk, real(aveuv(kk, :))   ! aveuv is real (engmain.F90:339)
          end do
          close(io)
       end if
       if (prm%slttype == 1) then
          open(newunit=io, file='weight_soln', action='write')
       else
          open(newunit=io, file='weight_refs', action='write')
       end if
       do kk = 1, engdiv
          if (prm%wgtslf == 0) then
             write(io, '(i5,e20.8)') kk, avediv(kk, 1)
          else
             write(io, '(i5,e20.8,e20.7)') kk, avediv(kk, 1), avediv(kk, 2)
          end if
       end do
       close(io)
       open(newunit=io, file='uvrange.tt', action='write')
       write(io, '(A)') ' species     minimum        maximum'
       do pti = 0, numslv
          if (maxuv(pti) < 1.0e5) then
             write(io, '(i5,2f15.5)') pti, minuv(pti), maxuv(pti)
          else
             write(io, '(i5,f15.5,g18.5)') pti, minuv(pti), maxuv(pti)
          end if
       end do
       close(io)
    end if
    if (engdiv == 1) then
       suffeng = '.tt'
    else
       j = division / 10
       kk = mod(division, 10)
       suffeng = '.' // numbers(j+1:j+1) // numbers(kk+1:kk+1)
    end if
    if (prm%slttype == 1) then
       open(newunit=io, file='engsln' // suffeng, form="FORMATTED", action='write')
    else
       open(newunit=io, file='engref' // suffeng, form="FORMATTED", action='write')
    end if
    write(io, '(A2,A13,A15,A5,A25,A25)') "# ", "bin_left", "bin_middle", "spec", "ratio", "dens"
    do iduv = 1, ermax
       call bin_info(iduv, leftbin, middlebin, invwt, pti)
       write(io, '(g15.7,g15.7,i5,g25.15,g25.15)') leftbin, middlebin, pti, edens(iduv), edens(iduv) * invwt
    end do
    close(io)
    if (prm%corrcal == 1) then
       if (prm%slttype == 1) then
          open(newunit=io, file='corsln' // suffeng, form="UNFORMATTED", action='write')
       else
          open(newunit=io, file='corref' // suffeng, form="UNFORMATTED", action='write')
       end if
       write(io) ecorr
       close(io)
    end if
    if (esmax > 0) then
       open(newunit=io, file='slfeng' // suffeng, form="FORMATTED", action='write')
       do iduv = 1, esmax
          write(io, '(g15.7,g25.15)') selfenergy_coord(iduv), eself(iduv)
       end do
       close(io)
    end if
  end subroutine engstore
end program erdst_host
