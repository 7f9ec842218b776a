! tce_t_mod.F90 -- ISO_C_BINDING interface of libtce_t.so for Fortran callers.
!
! The per-call kernel surface (sd_t_s1_N_cuda, sd_t_d1_N_cuda, sd_t_d2_N_cuda,
! compute_en, dev_mem_s/d, dev_release, initmemmodule, finalizememmodule,
! check_device, device_init) keeps the reference's implicit-interface calls of
! src/tce/ccsd_t/ccsd_t_gpu.F, ccsd_t_singles_gpu.F and ccsd_t_doubles_gpu.F:
! those files link against libtce_t.so unchanged.  This module adds explicit
! interfaces for the whole-job entry points (one device-resident job replacing
! the tile-sextuple loop of ccsd_t_gpu.F) and the host helpers.
!
! All INTEGERs at the boundary are integer(c_long), matching NWChem's
! -fdefault-integer-8 build.
module tce_t_mi355x
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TCE_T_OK = 0
  integer(c_int), parameter :: TCE_T_FP64 = 0
  integer(c_int), parameter :: TCE_T_LAMBDA_ORBITAL = 0, TCE_T_LAMBDA_REFERENCE_ORDER = 1
  integer(c_int), parameter :: TCE_T_TILING_REFERENCE = 0, TCE_T_TILING_SUBTILE = 1, &
                               TCE_T_TILING_ORBITAL = 2

  ! Mirror of tce_t_tiles: the tce.fh / MA arrays the (T) path reads.
  type, bind(C) :: tce_t_tiles
    integer(c_long) :: noab, nvab
    type(c_ptr) :: spin, sym, range, offset, alpha
    type(c_ptr) :: evl
    integer(c_int) :: restricted
  end type tce_t_tiles

  ! Orbital-form integrals (intorb): alpha tiles and k_b2am.
  type, bind(C) :: tce_t_intorb
    integer(c_long) :: noa, nva
    type(c_ptr) :: b2am, spin_alpha, sym_alpha, range_alpha
    integer(c_long) :: irrep_v
  end type tce_t_intorb

  interface
    integer(c_int) function tce_t_init(device) bind(C, name="tce_t_init")
      import :: c_int
      integer(c_int), value :: device
    end function

    integer(c_int) function tce_t_job_create(tiles, job) bind(C, name="tce_t_job_create")
      import :: c_int, c_ptr, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      type(c_ptr), intent(out) :: job
    end function

    integer(c_long) function tce_t_job_num_tasks(job) bind(C, name="tce_t_job_num_tasks")
      import :: c_long, c_ptr
      type(c_ptr), value :: job
    end function

    integer(c_long) function tce_t_job_device_bytes(job) bind(C, name="tce_t_job_device_bytes")
      import :: c_long, c_ptr
      type(c_ptr), value :: job
    end function

    ! d_t1/k_t1_offset, d_t2/k_t2_offset, d_v2/k_v2_offset of ccsd_t_gpu.F
    integer(c_int) function tce_t_job_upload(job, t1, k_t1_offset, t2, k_t2_offset, &
                                              v2, k_v2_offset) bind(C, name="tce_t_job_upload")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      real(c_double), intent(in) :: t1(*), t2(*), v2(*)
      integer(c_long), intent(in) :: k_t1_offset(*), k_t2_offset(*), k_v2_offset(*)
    end function

    ! energy1 = E[T], energy2 = E(T)
    integer(c_int) function tce_t_job_run(job, precision, energy1, energy2) &
        bind(C, name="tce_t_job_run")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: job
      integer(c_int), value :: precision
      real(c_double), intent(out) :: energy1, energy2
    end function

    ! ccsd_t_restart.F: blocks of one outer virtual tile; begin and table
    ! are what the caller keeps in the rtdb (tce:ccsd_t_restart_begin,
    ! tce:restart_triples_table); max_tiles <= 0 runs every remaining block
    integer(c_int) function tce_t_job_run_restart(job, precision, begin, table, max_tiles, &
        t_energy) bind(C, name="tce_t_job_run_restart")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      integer(c_int), value :: precision
      integer(c_long), intent(inout) :: begin
      real(c_double), intent(inout) :: table(*)
      integer(c_long), value :: max_tiles
      real(c_double), intent(out) :: t_energy
    end function

    ! kernel-task tiling: TCE_T_TILING_REFERENCE, _SUBTILE (default), _ORBITAL
    integer(c_int) function tce_t_job_create_tiling(tiles, tiling, job) &
        bind(C, name="tce_t_job_create_tiling")
      import :: c_int, c_ptr, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_int), value :: tiling
      type(c_ptr), intent(out) :: job
    end function

    ! ccsd_t_singles_l / ccsd_t_doubles_l (toggle 2) on the resident tensors:
    ! adds t3s to singles and t3d to doubles (triplesx(h3,h2,h1,p6,p5,p4) of
    ! the task with these 1-based tiles), for cr_ccsd_t, lambda_ccsd_t, emb2
    integer(c_int) function tce_t_job_triples_tiles(job, precision, h1b, h2b, h3b, p4b, p5b, &
        p6b, singles, doubles) bind(C, name="tce_t_job_triples_tiles")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      integer(c_int), value :: precision
      integer(c_long), value :: h1b, h2b, h3b, p4b, p5b, p6b
      real(c_double), intent(inout) :: singles(*), doubles(*)
    end function
    ! the memory-sliced drivers of ccsd_t_6dts.F (ccsd_t_singles_l_6dts,
    ! ccsd_t_doubles_l_6dts): one p4 x p5 slice of the task's tiles
    integer(c_int) function tce_t_job_triples_tiles_6dts(job, precision, h1b, h2b, h3b, p4b, &
        p5b, p6b, istart, istop, jstart, jstop, maxp4, maxp5, singles, doubles) &
        bind(C, name="tce_t_job_triples_tiles_6dts")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      integer(c_int), value :: precision
      integer(c_long), value :: h1b, h2b, h3b, p4b, p5b, p6b
      integer(c_long), value :: istart, istop, jstart, jstop, maxp4, maxp5
      real(c_double), intent(inout) :: singles(*), doubles(*)
    end function

    ! the same tiles left in device memory (device addresses; valid until the
    ! job's next tile, Lambda or upload call)
    integer(c_int) function tce_t_job_triples_tiles_device(job, precision, h1b, h2b, h3b, &
        p4b, p5b, p6b, singles, doubles, elements) bind(C, name="tce_t_job_triples_tiles_device")
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: job
      integer(c_int), value :: precision
      integer(c_long), value :: h1b, h2b, h3b, p4b, p5b, p6b
      type(c_ptr), intent(out) :: singles, doubles
      integer(c_long), intent(out) :: elements
    end function

    integer(c_int) function tce_t_job_task_energies(job, out, ntasks) &
        bind(C, name="tce_t_job_task_energies")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      real(c_double), intent(out) :: out(*)
      integer(c_long), value :: ntasks
    end function

    ! Lambda-CCSD(T) (lambda_ccsd_t.F): the left-hand job from d_y1/k_l1_offset,
    ! d_y2/k_l2_offset and the right-hand job's d_v2/k_v2_offset
    integer(c_int) function tce_t_job_upload_lambda(job, y1, k_y1_offset, y2, k_y2_offset, &
                                                     v2, k_v2_offset) &
        bind(C, name="tce_t_job_upload_lambda")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      real(c_double), intent(in) :: y1(*), y2(*), v2(*)
      integer(c_long), intent(in) :: k_y1_offset(*), k_y2_offset(*), k_v2_offset(*)
    end function

    ! energy1, energy2 of lambda_ccsd_t from the right-hand (t1, t2) and
    ! left-hand (Lambda) jobs and d_f1/k_f1_offset; task_energies(2, ntasks)
    integer(c_int) function tce_t_lambda_ccsd_t(right, left, f1, k_f1_offset, pairing, &
                                                 precision, energy1, energy2, task_energies, &
                                                 ntasks) bind(C, name="tce_t_lambda_ccsd_t")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: right, left
      real(c_double), intent(in) :: f1(*)
      integer(c_long), intent(in) :: k_f1_offset(*)
      integer(c_int), value :: pairing, precision
      real(c_double), intent(out) :: energy1, energy2
      real(c_double), intent(out) :: task_energies(*)
      integer(c_long), value :: ntasks
    end function

    ! A CR-CCSD(T) job: d_t1, d_t2 and the two intermediates cr_ccsd_t_N
    ! builds with toggle 1 (d_i1_1/k_i1_offset_1, d_i1_2/k_i1_offset_2)
    integer(c_int) function tce_t_job_upload_cr(job, t1, k_t1_offset, t2, k_t2_offset, &
                                                i1_1, k_i1_offset_1, i1_2, k_i1_offset_2) &
        bind(C, name="tce_t_job_upload_cr")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: job
      real(c_double), intent(in) :: t1(*), t2(*), i1_1(*), i1_2(*)
      integer(c_long), intent(in) :: k_t1_offset(*), k_t2_offset(*)
      integer(c_long), intent(in) :: k_i1_offset_1(*), k_i1_offset_2(*)
    end function

    ! energy1, energy2 of cr_ccsd_t from a (T) job and a CR job; terms(5) =
    ! num1, num2, den0, den1, den2
    integer(c_int) function tce_t_cr_ccsd_t(triples, cr, precision, energy1, energy2, terms) &
        bind(C, name="tce_t_cr_ccsd_t")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: triples, cr
      integer(c_int), value :: precision
      real(c_double), intent(out) :: energy1, energy2
      real(c_double), intent(out) :: terms(*)
    end function

    integer(c_int) function tce_t_job_destroy(job) bind(C, name="tce_t_job_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: job
    end function

    integer(c_long) function tce_t_tile_ranges(n, tilesize, ranges) bind(C, name="tce_t_tile_ranges")
      import :: c_long
      integer(c_long), value :: n, tilesize
      integer(c_long), intent(out) :: ranges(*)
    end function

    integer(c_long) function tce_t_t1_offset(tiles, hash, size) bind(C, name="tce_t_t1_offset")
      import :: c_long, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    integer(c_long) function tce_t_t2_offset(tiles, hash, size) bind(C, name="tce_t_t2_offset")
      import :: c_long, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    integer(c_long) function tce_t_v2_offset(tiles, hash, size) bind(C, name="tce_t_v2_offset")
      import :: c_long, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    ! tce_hash.F: offset of key in a block table, or -1
    integer(c_long) function tce_t_hash(hash, key) bind(C, name="tce_t_hash")
      import :: c_long
      integer(c_long), intent(in) :: hash(*)
      integer(c_long), value :: key
    end function

    ! tce_y1_offset.F, tce_y2_offset.F, tce_mo1e_offset.F
    integer(c_long) function tce_t_y1_offset(tiles, hash, size) bind(C, name="tce_t_y1_offset")
      import :: c_long, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    integer(c_long) function tce_t_y2_offset(tiles, hash, size) bind(C, name="tce_t_y2_offset")
      import :: c_long, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    integer(c_long) function tce_t_f1_offset(tiles, hash, size) bind(C, name="tce_t_f1_offset")
      import :: c_long, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    ! k_v2_alpha_offset of tce_mo2e_offset_intorb.F (length query: pass a
    ! one-element hash and size, then call again with the full table)
    integer(c_long) function tce_t_v2_alpha_offset(io, idiv2e, hash, size) &
        bind(C, name="tce_t_v2_alpha_offset")
      import :: c_long, tce_t_intorb
      type(tce_t_intorb), intent(in) :: io
      integer(c_long), value :: idiv2e
      integer(c_long), intent(out) :: hash(*)
      integer(c_long), intent(out) :: size
    end function

    ! get_hash_block_i over every block of v2_hash: the spin-orbital V2
    ! blocks tce_t_job_upload reads, from the orbital-form d_v2orb
    integer(c_int) function tce_t_v2_from_intorb(tiles, io, v2orb, v2_alpha_offset, v2_hash, v2) &
        bind(C, name="tce_t_v2_from_intorb")
      import :: c_int, c_long, c_double, tce_t_tiles, tce_t_intorb
      type(tce_t_tiles), intent(in) :: tiles
      type(tce_t_intorb), intent(in) :: io
      real(c_double), intent(in) :: v2orb(*)
      integer(c_long), intent(in) :: v2_alpha_offset(*)
      integer(c_long), intent(in) :: v2_hash(*)
      real(c_double), intent(out) :: v2(*)
    end function

    integer(c_int) function tce_t_synth_evl(seed, no, nv, evl) bind(C, name="tce_t_synth_evl")
      import :: c_int, c_long, c_long_long, c_double
      integer(c_long_long), value :: seed
      integer(c_long), value :: no, nv
      real(c_double), intent(out) :: evl(*)
    end function

    integer(c_int) function tce_t_synth_blocks(tiles, seed, t1, t1h, t2, t2h, v2, v2h) &
        bind(C, name="tce_t_synth_blocks")
      import :: c_int, c_long, c_long_long, c_double, tce_t_tiles
      type(tce_t_tiles), intent(in) :: tiles
      integer(c_long_long), value :: seed
      real(c_double), intent(out) :: t1(*), t2(*), v2(*)
      integer(c_long), intent(in) :: t1h(*), t2h(*), v2h(*)
    end function
  end interface
end module tce_t_mi355x
