! tce_t_fdemo.F90 -- a Fortran caller of the whole-job API: builds the tile
! model the way tce_tile.F does (holes first, then particles; C1, one spin),
! fills T1/T2/V2 canonical blocks with the seeded generator, uploads them with
! their hash tables and prints E[T], E(T); then E(T) by restart blocks, and
! E[T], E(T) once more from the task tiles the library hands back, summed
! here as the tiles' other consumers (cr_ccsd_t.F, lambda_ccsd_t.F) do;
! then the Lambda-CCSD(T) energies (lambda_ccsd_t.F) of a second seeded
! amplitude set taken as the Lambda amplitudes; then CR-CCSD(T)
! (cr_ccsd_t.F) with the bare integrals as its two moment intermediates.
! Usage: tce_t_fdemo [no nv tile_o tile_v]   (default 8 32 8 16)
!        tce_t_fdemo intorb   (host only: the orbital-form integral input of a
!        closed-shell tiling through tce_t_v2_alpha_offset and
!        tce_t_v2_from_intorb, printed as exact checksums)
program tce_t_fdemo
  use iso_c_binding
  use tce_t_mi355x
  implicit none
  integer(c_long) :: no, nv, to, tv, noab, nvab, n, i, l1, l2, lv, s1, s2, sv
  integer(c_long), allocatable, target :: range(:), offset(:), spin(:), sym(:), alpha(:)
  integer(c_long), allocatable :: h1(:), h2(:), hv(:)
  real(c_double), allocatable, target :: evl(:)
  real(c_double), allocatable :: t1(:), t2(:), v2(:)
  type(tce_t_tiles) :: tiles
  type(c_ptr) :: job
  real(c_double) :: e1, e2, et
  real(c_double), allocatable :: table(:)
  integer(c_long) :: begin
  integer(c_long) :: p4b, p5b, p6b, h1b, h2b, h3b, a4, a5, a6, i1, i2, i3, sz
  real(c_double), allocatable :: singles(:), doubles(:)
  real(c_double) :: factor, den, ta, tb
  integer(c_int) :: rc
  character(len=32) :: arg
  integer(c_long_long), parameter :: seed = 1234_c_long_long

  no = 8; nv = 32; to = 8; tv = 16
  if (command_argument_count() == 1) then
    call get_command_argument(1, arg)
    if (trim(arg) == 'intorb') then
      call intorb_demo()
      stop
    end if
  end if
  if (command_argument_count() >= 4) then
    call get_command_argument(1, arg); read(arg, *) no
    call get_command_argument(2, arg); read(arg, *) nv
    call get_command_argument(3, arg); read(arg, *) to
    call get_command_argument(4, arg); read(arg, *) tv
  end if
  allocate(range(no + nv))
  noab = tce_t_tile_ranges(no, to, range)
  nvab = tce_t_tile_ranges(nv, tv, range(noab + 1:))
  n = noab + nvab
  allocate(offset(n), spin(n), sym(n), alpha(n), evl(no + nv))
  offset(1) = 0
  do i = 2, n
    offset(i) = offset(i - 1) + range(i - 1)
  end do
  spin = 1; sym = 0
  do i = 1, n
    alpha(i) = i
  end do
  rc = tce_t_synth_evl(seed, no, nv, evl)
  tiles = tce_t_tiles(noab, nvab, c_loc(spin), c_loc(sym), c_loc(range), c_loc(offset), &
                      c_loc(alpha), c_loc(evl), 0)
  allocate(h1(2 * noab * nvab + 1), h2(2 * (noab * nvab)**2 + 1), hv(2 * n**4 + 1))
  l1 = tce_t_t1_offset(tiles, h1, s1)
  l2 = tce_t_t2_offset(tiles, h2, s2)
  lv = tce_t_v2_offset(tiles, hv, sv)
  allocate(t1(s1), t2(s2), v2(sv))
  rc = tce_t_synth_blocks(tiles, seed, t1, h1, t2, h2, v2, hv)
  if (rc /= TCE_T_OK) stop 2
  rc = tce_t_init(0_c_int)
  if (rc /= TCE_T_OK) stop 3
  rc = tce_t_job_create(tiles, job)
  if (rc /= TCE_T_OK) stop 4
  rc = tce_t_job_upload(job, t1, h1, t2, h2, v2, hv)
  if (rc /= TCE_T_OK) stop 5
  rc = tce_t_job_run(job, TCE_T_FP64, e1, e2)
  if (rc /= TCE_T_OK) stop 6
  write(*, '(A,I0,A,I0,A,ES24.16,A,ES24.16)') 'tasks=', tce_t_job_num_tasks(job), &
      ' blocks(t2)=', l2, ' E[T]=', e1, ' E(T)=', e2
  ! ccsd_t_restart.F pattern: one outer virtual tile per call, the state
  ! (begin, table) carried between calls as the rtdb would keep it
  allocate(table(nvab))
  table = 0.0d0
  begin = 1
  do while (begin <= nvab)
    rc = tce_t_job_run_restart(job, TCE_T_FP64, begin, table, 1_c_long, et)
    if (rc /= TCE_T_OK) stop 7
  end do
  write(*, '(A,I0,A,ES24.16)') 'restart blocks=', nvab, ' E(T)=', et
  ! the task loop of ccsd_t_gpu.F (C1, one spin: every sextuple is a task)
  ! with the tiles from tce_t_job_triples_tiles and the energy of
  ! ccsd_t_gpu.F:153-200 evaluated by the caller
  ta = 0.0d0; tb = 0.0d0
  do p4b = noab + 1, n
  do p5b = p4b, n
  do p6b = p5b, n
  do h1b = 1, noab
  do h2b = h1b, noab
  do h3b = h2b, noab
    sz = range(p4b) * range(p5b) * range(p6b) * range(h1b) * range(h2b) * range(h3b)
    allocate(singles(sz), doubles(sz))
    singles = 0.0d0; doubles = 0.0d0
    rc = tce_t_job_triples_tiles(job, TCE_T_FP64, h1b, h2b, h3b, p4b, p5b, p6b, singles, doubles)
    if (rc /= TCE_T_OK) stop 8
    factor = 1.0d0
    if (p4b == p5b .and. p5b == p6b) then
      factor = factor / 6.0d0
    else if (p4b == p5b .or. p5b == p6b) then
      factor = factor / 2.0d0
    end if
    if (h1b == h2b .and. h2b == h3b) then
      factor = factor / 6.0d0
    else if (h1b == h2b .or. h2b == h3b) then
      factor = factor / 2.0d0
    end if
    i = 0
    do a4 = 1, range(p4b)
    do a5 = 1, range(p5b)
    do a6 = 1, range(p6b)
    do i1 = 1, range(h1b)
    do i2 = 1, range(h2b)
    do i3 = 1, range(h3b)
      i = i + 1
      den = evl(offset(h1b) + i1) + evl(offset(h2b) + i2) + evl(offset(h3b) + i3) &
          - evl(offset(p4b) + a4) - evl(offset(p5b) + a5) - evl(offset(p6b) + a6)
      ta = ta + factor * doubles(i) * doubles(i) / den
      tb = tb + factor * doubles(i) * (doubles(i) + singles(i)) / den
    end do
    end do
    end do
    end do
    end do
    end do
    deallocate(singles, doubles)
  end do
  end do
  end do
  end do
  end do
  end do
  write(*, '(A,ES24.16,A,ES24.16)') 'tiles E[T]=', ta, ' E(T)=', tb
  call lambda_demo()
  call cr_demo()
  rc = tce_t_job_destroy(job)

contains

  ! The Lambda amplitudes y1, y2 laid out as tce_y1_offset / tce_y2_offset
  ! keep them (blocks [h][p], [h][h'][p][p']), taken from the t-layout blocks
  ! of the generator's seed 4321; f1 over the tce_mo1e_offset table as
  ! 0.05 sin(k); the left-hand job uploaded from them, and the energies of
  ! lambda_ccsd_t with the job above as the right-hand side.
  subroutine lambda_demo()
    integer(c_long), allocatable :: hy1(:), hy2(:), hf(:)
    real(c_double), allocatable :: u1(:), u2(:), uv(:), y1(:), y2(:), f1(:), te(:)
    integer(c_long) :: ly1, ly2, lf, sy1, sy2, sf, k, key, off, tof, h, p, q1, q2, g3, g4
    integer(c_long) :: ra, rb, ri, rj, a, b, ii, jj, nt
    type(c_ptr) :: left
    real(c_double) :: l1, l2
    allocate(hy1(2 * noab * nvab + 1), hy2(2 * (noab * nvab)**2 + 1), hf(2 * n * n + 1))
    ly1 = tce_t_y1_offset(tiles, hy1, sy1)
    ly2 = tce_t_y2_offset(tiles, hy2, sy2)
    lf = tce_t_f1_offset(tiles, hf, sf)
    allocate(u1(s1), u2(s2), uv(sv), y1(sy1), y2(sy2), f1(sf))
    rc = tce_t_synth_blocks(tiles, 4321_c_long_long, u1, h1, u2, h2, uv, hv)
    if (rc /= TCE_T_OK) stop 9
    do k = 1, ly1  ! y(h p) block [h][p] from the t1 block [p][h]
      key = hy1(1 + k)
      off = hy1(1 + ly1 + k)
      h = key / nvab + 1
      p = mod(key, nvab) + noab + 1
      tof = tce_t_hash(h1, h - 1 + noab * (p - noab - 1))
      ri = range(h)
      ra = range(p)
      do ii = 0, ri - 1
        do a = 0, ra - 1
          y1(off + ii * ra + a + 1) = u1(tof + a * ri + ii + 1)
        end do
      end do
    end do
    do k = 1, ly2  ! y(h3 h4 p1 p2) block [h3][h4][p1][p2] from the t2 block [p1][p2][h3][h4]
      key = hy2(1 + k)
      off = hy2(1 + ly2 + k)
      q2 = mod(key, nvab) + noab + 1
      key = key / nvab
      q1 = mod(key, nvab) + noab + 1
      key = key / nvab
      g4 = mod(key, noab) + 1
      g3 = key / noab + 1
      tof = tce_t_hash(h2, g4 - 1 + noab * (g3 - 1 + noab * (q2 - noab - 1 + nvab * (q1 - noab - 1))))
      ra = range(q1)
      rb = range(q2)
      ri = range(g3)
      rj = range(g4)
      do ii = 0, ri - 1
        do jj = 0, rj - 1
          do a = 0, ra - 1
            do b = 0, rb - 1
              y2(off + ((ii * rj + jj) * ra + a) * rb + b + 1) = &
                  u2(tof + ((a * rb + b) * ri + ii) * rj + jj + 1)
            end do
          end do
        end do
      end do
    end do
    do k = 1, sf
      f1(k) = 0.05d0 * sin(dble(k))
    end do
    rc = tce_t_job_create(tiles, left)
    if (rc /= TCE_T_OK) stop 10
    rc = tce_t_job_upload_lambda(left, y1, hy1, y2, hy2, v2, hv)
    if (rc /= TCE_T_OK) stop 11
    nt = tce_t_job_num_tasks(job)
    allocate(te(2 * nt))
    rc = tce_t_lambda_ccsd_t(job, left, f1, hf, TCE_T_LAMBDA_ORBITAL, TCE_T_FP64, l1, l2, te, nt)
    if (rc /= TCE_T_OK) stop 12
    write(*, '(A,ES24.16,A,ES24.16)') 'lambda energy1=', l1, ' energy2=', l2
    rc = tce_t_job_destroy(left)
  end subroutine lambda_demo

  ! CR-CCSD(T) with the bare integrals as cr_ccsd_t_N's two intermediates:
  ! i1(h11 p4 h1 h2) = v(h11 p4 h1 h2) under the table of
  ! OFFSET_cr_ccsd_t_N_1_1 (blocks [p4][h11][h1][h2]) and i1(p4 p5 h1 p12) =
  ! v(p4 p5 h1 p12) under OFFSET_cr_ccsd_t_N_2_1 (blocks [p4][p5][h1][p12]),
  ! copied from the v2 blocks (every block is allowed: one spin, one irrep).
  ! The moment sums num1, num2 are then E[T], E(T) of the job above.
  subroutine cr_demo()
    integer(c_long), allocatable :: hi1(:), hi2(:)
    real(c_double), allocatable :: i1a(:), i1b(:)
    integer(c_long) :: la, lb, sa, sb, k, off, vo, m, a, ii, jj, nb
    integer(c_long) :: p4, p5, h11, hh1, hh2, p12, ra, rm, ri, rj
    type(c_ptr) :: crj
    real(c_double) :: c1, c2, terms(5)
    la = nvab * noab * noab * (noab + 1) / 2
    lb = nvab * (nvab + 1) / 2 * noab * nvab
    allocate(hi1(2 * la + 1), hi2(2 * lb + 1))
    hi1(1) = la
    hi2(1) = lb
    k = 0
    sa = 0
    do p4 = noab + 1, n
      do h11 = 1, noab
        do hh1 = 1, noab
          do hh2 = hh1, noab
            k = k + 1
            hi1(1 + k) = hh2 - 1 + noab * (hh1 - 1 + noab * (h11 - 1 + noab * (p4 - noab - 1)))
            hi1(1 + la + k) = sa
            sa = sa + range(p4) * range(h11) * range(hh1) * range(hh2)
          end do
        end do
      end do
    end do
    k = 0
    sb = 0
    do p4 = noab + 1, n
      do p5 = p4, n
        do hh1 = 1, noab
          do p12 = noab + 1, n
            k = k + 1
            hi2(1 + k) = p12 - noab - 1 + nvab * (hh1 - 1 + noab * (p5 - noab - 1 + nvab * (p4 - noab - 1)))
            hi2(1 + lb + k) = sb
            sb = sb + range(p4) * range(p5) * range(hh1) * range(p12)
          end do
        end do
      end do
    end do
    allocate(i1a(max(sa, 1_c_long)), i1b(max(sb, 1_c_long)))
    k = 0
    do p4 = noab + 1, n
      do h11 = 1, noab
        do hh1 = 1, noab
          do hh2 = hh1, noab
            k = k + 1
            off = hi1(1 + la + k)
            vo = tce_t_hash(hv, hh2 - 1 + n * (hh1 - 1 + n * (p4 - 1 + n * (h11 - 1))))
            ra = range(p4)
            rm = range(h11)
            ri = range(hh1)
            rj = range(hh2)
            do a = 0, ra - 1
              do m = 0, rm - 1
                do ii = 0, ri - 1
                  do jj = 0, rj - 1
                    i1a(off + ((a * rm + m) * ri + ii) * rj + jj + 1) = &
                        v2(vo + ((m * ra + a) * ri + ii) * rj + jj + 1)
                  end do
                end do
              end do
            end do
          end do
        end do
      end do
    end do
    k = 0
    do p4 = noab + 1, n
      do p5 = p4, n
        do hh1 = 1, noab
          do p12 = noab + 1, n
            k = k + 1
            off = hi2(1 + lb + k)
            vo = tce_t_hash(hv, p12 - 1 + n * (hh1 - 1 + n * (p5 - 1 + n * (p4 - 1))))
            nb = range(p4) * range(p5) * range(hh1) * range(p12)
            i1b(off + 1:off + nb) = v2(vo + 1:vo + nb)
          end do
        end do
      end do
    end do
    rc = tce_t_job_create(tiles, crj)
    if (rc /= TCE_T_OK) stop 13
    rc = tce_t_job_upload_cr(crj, t1, h1, t2, h2, i1a, hi1, i1b, hi2)
    if (rc /= TCE_T_OK) stop 14
    rc = tce_t_cr_ccsd_t(job, crj, TCE_T_FP64, c1, c2, terms)
    if (rc /= TCE_T_OK) stop 15
    write(*, '(A,5ES24.16)') 'cr terms=', terms
    write(*, '(A,ES24.16,A,ES24.16)') 'cr energy1=', c1, ' energy2=', c2
    rc = tce_t_job_destroy(crj)
  end subroutine cr_demo

  ! Closed-shell tiling as tce_tile.F lays it out with intorb: alpha holes
  ! (3, 2), beta holes, alpha particles (4, 5, 3), beta particles; k_b2am
  ! maps each tile to its alpha tile.  The orbital-form array holds exact
  ! binary fractions, so the checksums of the converted blocks are exact.
  subroutine intorb_demo()
    integer(c_long), parameter :: noa = 2, nva = 3, nt = 10, idiv2e = 2
    integer(c_long), target :: r(nt), o(nt), sp(nt), sy(nt), al(nt), b2am(nt)
    integer(c_long), target :: spa(noa + nva), sya(noa + nva), ra(noa + nva)
    real(c_double), target :: ev(34)   ! one energy per orbital (sum of r)
    integer(c_long), allocatable :: hv2(:)
    integer(c_long) :: ha(6 * (idiv2e + 2) + 1), words, osize, nv2, vsize, k
    real(c_double), allocatable :: orb(:), v(:)
    type(tce_t_tiles) :: ti
    type(tce_t_intorb) :: io
    real(c_double) :: total, check
    integer(c_int) :: rc2
    r = [3, 2, 3, 2, 4, 5, 3, 4, 5, 3]
    sp = [1, 1, 2, 2, 1, 1, 1, 2, 2, 2]
    sy = 0
    al = [1, 2, 1, 2, 5, 6, 7, 5, 6, 7]
    b2am = [1, 2, 1, 2, 3, 4, 5, 3, 4, 5]
    o(1) = 0
    do k = 2, nt
      o(k) = o(k - 1) + r(k - 1)
    end do
    do k = 1, 34
      ev(k) = -2.0d0 + 0.125d0 * k
    end do
    spa = 1; sya = 0; ra = [3, 2, 4, 5, 3]
    ti = tce_t_tiles(4_c_long, 6_c_long, c_loc(sp), c_loc(sy), c_loc(r), c_loc(o), c_loc(al), &
                     c_loc(ev), 1)
    io = tce_t_intorb(noa, nva, c_loc(b2am), c_loc(spa), c_loc(sya), c_loc(ra), 0_c_long)
    words = tce_t_v2_alpha_offset(io, idiv2e, ha, osize)
    if (words < 0) stop 11
    allocate(hv2(2 * nt**4 + 1))
    nv2 = tce_t_v2_offset(ti, hv2, vsize)
    allocate(orb(osize), v(vsize))
    do k = 1, osize
      orb(k) = dble(mod(k * 7919_c_long, 1000_c_long)) / 1024.0d0
    end do
    rc2 = tce_t_v2_from_intorb(ti, io, orb, ha, hv2, v)
    if (rc2 /= TCE_T_OK) stop 12
    total = 0.0d0; check = 0.0d0
    do k = 1, vsize
      total = total + v(k)
      check = check + v(k) * dble(mod(k, 17_c_long))
    end do
    write(*, '(a,i0,a,i0,a,es26.17e3,a,es26.17e3)') 'intorb blocks=', nv2, ' size=', vsize, &
        ' sum=', total, ' check=', check
  end subroutine intorb_demo
end program tce_t_fdemo
