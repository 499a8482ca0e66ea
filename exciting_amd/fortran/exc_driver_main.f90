! Test program for the Fortran driver: reads one k-point's inputs from a stream file,
! runs match and hamiltonsetup + overlapsetup (or hamiltonandoverlapsetup) through the GPU
! driver as seceqnfv.f90:73-77 would, and writes apwalm, H and O to a stream file.  With a third
! file of eigenpairs it also runs forcek_gpu and appends forceibs(3, natmtot), then
! gendmatmt_gpu for the same states and appends dm(wfsize, wfsize, natmtot) and genrhoir_gpu's
! rhoir(ngrtot) (and, with radial data, genrhomt_gpu's rhomt(lmmaxvr, nrmtmax, natmtot), then, if
! lmaxvr <= lmaxapw, rhovalk_gpu's rhomt of the same states by the direct route); with a fourth
! argument nev > 0 it then solves the eigenproblem (solvewithgpu, as seceqnfv.f90 calls
! solvewithlapack) and appends evalfv(nev) and evecfv(nmat, nev).  With a fifth file of radial
! data it first runs set_radial_gpu (match's boundary values from apwfr) and hmlint_gpu (hmlint +
! olprad), whose results replace the input's dmat and muffin-tin matrices; with a sixth file
! (bvec, V(r)) it runs gencfun_gpu and genveffig_gpu, whose Theta(G) and V(G) replace the input's.
!   exc_driver_main <input.bin> <output.bin> [<eigen.bin> | -] [nev] [<radial.bin> | -] [<scf.bin>]
program exc_driver_main
  use, intrinsic :: iso_c_binding
  use exc_lapw_c, only: exc_dims
  use exc_evsystem
  use exc_setup_gpu
  implicit none
  type(exc_dims) :: dims
  integer :: ngp, ngkmax, nmat, lmmax, naa, nbox, u, nlotot
  integer(c_int), allocatable :: natoms(:), nlorb(:), lorbl(:), idxlo(:), ivg(:), ivgig(:), igpig(:)
  real(8), allocatable :: rmt(:), atposc(:), dmat(:), oalo(:), ololo(:), ffacg(:), gpc(:), tpgpc(:, :), vgpc(:, :)
  complex(8), allocatable :: haa(:), loa(:), lolo(:), cfunig(:), veffig(:), sfacgp(:, :), apwalm(:, :, :, :)
  type(evsystem) :: system
  character(len=512) :: fin, fout, feig, fnev
  integer :: hdr(18), nst, u2, nev_solve
  complex(8), allocatable :: evec(:, :)
  real(8), allocatable :: evalfv(:), occ(:), forceibs(:, :), evs(:)
  complex(8), allocatable :: vecs(:, :), dm(:, :, :)
  real(8) :: wkpt
  integer :: ws, lmaxvr, nrmtmax
  character(len=512) :: frad, fscf
  real(8) :: bvec(3, 3)
  real(8), allocatable :: veffir(:), cfunir(:)
  integer(c_int), allocatable :: nrmt(:)
  real(8), allocatable :: spr(:, :), apwfr(:), lofr(:), veffmt(:), rhomt(:), rhoir(:), rfsht(:, :)
  complex(8), allocatable :: zbsht(:, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  feig = ""
  if (command_argument_count() >= 3) call get_command_argument(3, feig)
  nev_solve = 0
  if (command_argument_count() >= 4) then
    call get_command_argument(4, fnev)
    read(fnev, *) nev_solve
  end if
  open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
  read(u) hdr
  dims%natmtot = hdr(1); dims%nspecies = hdr(2); dims%lmaxapw = hdr(3); dims%lmaxmat = hdr(4)
  dims%apwordmax = hdr(5); dims%nlomax = hdr(6); dims%lolmmax = hdr(7); dims%maxnlo = hdr(8)
  dims%ngrtot = hdr(9); dims%ngvec = hdr(10); dims%intgv = hdr(11:16)
  ngp = hdr(17); ngkmax = hdr(18)
  read(u) dims%omega
  lmmax = (dims%lmaxapw + 1)**2
  naa = dims%apwordmax * (dims%lmaxmat + 1)**2
  nbox = product(dims%intgv(4:6) - dims%intgv(1:3) + 1)
  allocate(natoms(dims%nspecies), rmt(dims%nspecies), atposc(3 * dims%natmtot), nlorb(dims%nspecies))
  allocate(lorbl(dims%nlomax * dims%nspecies), idxlo(dims%lolmmax * dims%nlomax * dims%natmtot))
  allocate(dmat(dims%apwordmax**2 * (dims%lmaxapw + 1) * dims%natmtot))
  allocate(haa(naa * naa * dims%natmtot), loa(dims%maxnlo * naa * dims%natmtot))
  allocate(lolo(dims%maxnlo**2 * dims%natmtot), oalo(dims%apwordmax * dims%nlomax * dims%natmtot))
  allocate(ololo(dims%nlomax**2 * dims%natmtot), ivg(3 * dims%ngrtot), ivgig(nbox))
  allocate(cfunig(dims%ngrtot), veffig(dims%ngvec), ffacg(dims%ngvec * dims%nspecies))
  allocate(gpc(ngkmax), tpgpc(2, ngkmax), sfacgp(ngkmax, dims%natmtot), igpig(ngkmax), vgpc(3, ngkmax))
  read(u) natoms, rmt, atposc, nlorb, lorbl, idxlo, dmat, haa, loa, lolo, oalo, ololo
  read(u) ivg, ivgig, cfunig, veffig, ffacg
  read(u) gpc, tpgpc, sfacgp, igpig, vgpc
  close(u)

  call exc_gpu_init(dims, ngkmax, natoms, rmt, atposc, nlorb, lorbl, idxlo, dmat, haa, loa, lolo, &
                    oalo, ololo, ivg, ivgig, cfunig, veffig, ffacg)
  frad = "-"
  if (command_argument_count() >= 5) call get_command_argument(5, frad)
  if (trim(frad) /= "-") then
    ! radial data file: lmaxvr, nrmtmax, nrmt(nspecies), spr, apwfr, lofr, veffmt, and the
    ! spherical-harmonic transforms zbshtvr, rfshtvr (lmmaxvr x lmmaxvr) of genshtmat
    open(newunit=u, file=trim(frad), access="stream", form="unformatted", status="old")
    read(u) lmaxvr, nrmtmax
    allocate(nrmt(dims%nspecies), spr(nrmtmax, dims%nspecies))
    allocate(apwfr(nrmtmax * 2 * dims%apwordmax * (dims%lmaxapw + 1) * dims%natmtot))
    allocate(lofr(nrmtmax * 2 * max(dims%nlomax, 1) * dims%natmtot))
    allocate(veffmt((lmaxvr + 1)**2 * nrmtmax * dims%natmtot))
    allocate(zbsht((lmaxvr + 1)**2, (lmaxvr + 1)**2), rfsht((lmaxvr + 1)**2, (lmaxvr + 1)**2))
    read(u) nrmt, spr, apwfr, lofr, veffmt, zbsht, rfsht
    close(u)
    call set_radial_gpu(nrmt, spr, apwfr)
    call hmlint_gpu(lmaxvr, nrmt, spr, apwfr, lofr, veffmt, haa, loa, lolo, oalo, ololo)
  end if
  if (command_argument_count() >= 6) then
    ! interstitial file: bvec(3,3), veffir(ngrtot) -> Theta(G), Theta(r), V(G) on the GPU
    call get_command_argument(6, fscf)
    open(newunit=u, file=trim(fscf), access="stream", form="unformatted", status="old")
    allocate(veffir(dims%ngrtot), cfunir(dims%ngrtot))
    read(u) bvec, veffir
    close(u)
    call gencfun_gpu(bvec, cfunig, cfunir)
    call genveffig_gpu(veffir, veffig)
  end if
  nlotot = exc_gpu_nlotot()
  nmat = ngp + nlotot

  allocate(apwalm(ngkmax, dims%apwordmax, lmmax, dims%natmtot))
  apwalm = (0.d0, 0.d0)
  call match(ngp, gpc, tpgpc, sfacgp, apwalm)
  call newsystem(system, .false., nmat)
  ! EXC_DRIVER_SETUP selects the calls: "split" (default) hamiltonsetup then overlapsetup as
  ! seceqnfv.f90:76-77 makes them, "fused" hamiltonandoverlapsetup; "copy" the split calls
  ! with apwalm copied from the host again (exc_resident_apwalm = .false.)
  call get_environment_variable("EXC_DRIVER_SETUP", fnev)
  if (trim(fnev) == "fused") then
    call hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
  else
    if (trim(fnev) == "copy") exc_resident_apwalm = .false.
    call hamiltonsetup(system, ngp, apwalm, igpig, vgpc)
    call overlapsetup(system, ngp, apwalm, igpig, vgpc)
  end if

  open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(u) nmat
  write(u) apwalm
  write(u) system%hamilton%za
  write(u) system%overlap%za
  if (len_trim(feig) > 0 .and. trim(feig) /= "-") then
    open(newunit=u2, file=trim(feig), access="stream", form="unformatted", status="old")
    read(u2) nst, wkpt
    allocate(evec(nmat, nst), evalfv(nst), occ(nst), forceibs(3, dims%natmtot))
    read(u2) evec, evalfv, occ
    close(u2)
    forceibs = 0.d0
    call forcek_gpu(ngp, apwalm, igpig, vgpc, sfacgp, nmat, nst, evec, evalfv, occ, wkpt, forceibs)
    write(u) forceibs
    ws = exc_gpu_dmat_wfsize()
    allocate(dm(ws, ws, dims%natmtot))
    dm = (0.d0, 0.d0)
    call gendmatmt_gpu(ngp, apwalm, nmat, nst, evec, occ, wkpt, dm)
    write(u) dm
    allocate(rhoir(dims%ngrtot))
    rhoir = 0.d0
    call genrhoir_gpu(ngp, igpig, nmat, nst, evec, occ, wkpt, 1.d-8, rhoir)
    write(u) rhoir
    if (trim(frad) /= "-") then
      ! with radial data: the muffin-tin density of these states
      allocate(rhomt((lmaxvr + 1)**2 * nrmtmax * dims%natmtot))
      call genrhomt_gpu(dm, rhomt)
      write(u) rhomt
      if (lmaxvr <= dims%lmaxapw) then
        ! and by the direct route (rhovalk, lradstep 1), where wavefmt allows lmaxvr
        rhomt = 0.d0
        call rhovalk_gpu(1, zbsht, rfsht, ngp, apwalm, nmat, nst, evec, occ, wkpt, 1.d-8, rhomt)
        write(u) rhomt
      end if
    end if
  end if
  if (nev_solve > 0) then
    allocate(evs(nev_solve), vecs(nmat, nev_solve))
    call solvewithgpu(system, nev_solve, vecs, evs)
    write(u) evs
    write(u) vecs
  end if
  close(u)
  call deletesystem(system)
  call exc_gpu_finalize()
end program exc_driver_main
