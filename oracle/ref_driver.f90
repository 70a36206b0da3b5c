! Fixture driver for the SMMP Fortran reference (test infrastructure only).
!
! Linked against the reference's own objects, built from the sources where
! they lie (oracle/Makefile); nothing of the reference is copied here.  Each
! mode prints '@name type count' headers followed by one value per line so
! tests/golden/make_golden.py can parse the output.
!
! Modes (first command-line argument):
!   mt     seed n                                  MT19937 stream (twister.f)
!   topo   libdir seqfile varfile itysol           topology, lists, tables
!   energy libdir seqfile varfile itysol nconf seed  random conformations
!   canon  libdir seqfile varfile itysol seed temp nsweep  Metropolis run
!   metro  libdir seqfile varfile itysol seed temp nsweep upch bgsprob rndord wfile
!          Metropolis run with /updchois/ set (upchswitch, bgsprob, rndord
!          0/1) and the weight can_weight (wfile '-') or b(k)x + alpha(k)
!          from wfile ('kmin kmax ebin' then one 'b alpha' line per bin);
!          prints every sweep's eol and acz, then /updstats/
!   partem libdir seqfile itysol seed nrep nequi nswp [switch]  partem_s run
!   muca   libdir seqfile itysol seed nsweep nup temp kmin kmax ebin
!   helix  libdir seqfile itysol seed nconf         helix() of drawn conformations
!   obs    libdir seqfile itysol seed nconf prefix [gbfile boxsize]
!          hbond, interhbond, zimmer and outpdb of drawn conformations
!   esolan libdir seqfile radfile seed nconf       analytic solvent (itysol -1):
!          energies of the start and of random conformations, and the
!          twister position after each (esolan.f draws on degenerate geometry);
!          radfile ('-' or 'index radians' lines) sets start variables
!   grad   libdir seqfile varfile itysol nconf seed epsd
!          gradient() (gradient.f, opeshe.f) at the start and at nconf
!          random conformations, epsd 0/1 the distance-dependent dielectric
!   minim  libdir seqfile varfile itysol imin maxit eps nconf seed
!          the minimisation of minim(imin, maxit, eps) (minim.f:35-160 with the
!          reference's minqsn.f / mincjg.f and move) from the start and from
!          nconf random conformations: start, final variables, final
!          energies, function calls
!   rmsd   libdir seqfile varfile itysol seed
!          rmsdfun (rmsdfun.f) of the start against a reference made from
!          it (turned, shifted, perturbed by the twister; a few atoms
!          without a reference), for isl 0, 1, 2 and two residue ranges
!   multi  libdir seqfile itysol gbfile boxsize seed nconf temp nequi nsweep
!          several chains: gbfile holds 'molecule k value' lines for gbpr;
!          topology, energies of random conformations with the
!          intermolecular terms (eninteract.f), then a Metropolis run with
!          the rigid-body trials of metropolis.f:94-168 (nequi unrecorded
!          sweeps, then the state and nsweep recorded sweeps)
! Tables of a multicanonical weight b(k) x + alpha(k) for the 'metro' mode.
! The reference keeps its own tables private to module multicanonical and
! only fills them inside mulcan_par / mulcan_sim, which stop under this
! compiler before their sweeps (an internal write of two records into one,
! mulcan_par_mod.f90:105 and the same format in mulcan_sim), so the driver
! passes this function, the formula of muca_weight2
! (mulcan_par_mod.f90:398-412), to the reference's own metropolis.
module drv_weights
  implicit none
  double precision, allocatable :: wb(:), wal(:)
  double precision :: webin, wxmin, wxmax
contains
  double precision function drv_muca2(x)
    double precision, intent(in) :: x
    integer :: k
    k = int(min(wxmax, max(wxmin, x/webin + sign(0.5d0, x))))
    drv_muca2 = wb(k)*x + wal(k)
  end function
end module drv_weights

program ref_driver
  use multicanonical
  use drv_weights
  include 'INCL.H'
  double precision energy, grnd, addang
  external can_weight
  character(len=256) :: mode, arg
  character(len=80) :: libdir, seqf, varf, radf
  character(len=4) :: grpn, grpc
  integer :: iseed, n, i, j, k, nconf, nsw, isol, nrep, nequi, nswp
  integer :: kmn, kmx, nup, iswt, ios, nhel, mhel, nbet, mbet
  double precision :: vrad
  double precision :: e, eolx, acz, tmp, ebw
  double precision :: ggasc
  parameter (ggasc = 1.98773d-3)
  integer :: ncalls, nacalls, kk
  common/updstats/ncalls(5),nacalls(5)
  character(len=256) :: wfile
  integer :: iepsd, nvta, imin, maxit
  double precision :: epsmin
  double precision :: rmsdfun, xrf(10000), yrf(10000), zrf(10000), ang(3), rot(3,3), px, py, pz
  integer :: ixat(mxat), nat, isl, ir1s(2), ir2s(2), q

  call get_command_argument(1, mode)
  grpn = 'nh2 '
  grpc = 'cooh'

  select case (trim(mode))
  case ('mt')
    call iarg(2, iseed)
    call iarg(3, n)
    call sgrnd(iseed)
    write(*,'(a,i0)') '@mt double ', n
    do i = 1, n
      write(*,'(es25.17)') grnd()
    end do

  case ('topo')
    call setup(2, .true.)
    call dump_params()
    call dump_topo()
    e = energy()
    call dump_energy(e)
    call dump_coords()

  case ('energy')
    call setup(2, .true.)
    call iarg(6, nconf)
    call iarg(7, iseed)
    call sgrnd(iseed)
    write(*,'(a,i0)') '@nconf int 1'
    write(*,'(i0)') nconf
    do k = 1, nconf
      do i = 1, nvr
        j = idvr(i)
        vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
      end do
      e = energy()
      write(*,'(a,i0)') '@vlvr double ', nvr
      do i = 1, nvr
        write(*,'(es25.17)') vlvr(idvr(i))
      end do
      call dump_energy(e)
      if (k .le. 2) call dump_coords()
    end do

  case ('helix')
    ! helix(nhel,mhel,nbet,mbet) of conformations whose free variables are
    ! drawn near the helical (-70,-37) or sheet (-150,150) limits or at random
    call setup(2, .false.)
    call iarg(5, iseed)
    call iarg(6, nconf)
    call sgrnd(iseed)
    write(*,'(a,i0)') '@nconf int 1'
    write(*,'(i0)') nconf
    do k = 1, nconf
      do i = 1, nvr
        j = idvr(i)
        tmp = grnd()
        if (tmp .lt. 0.35d0) then
          vlvr(j) = (-70.0d0 + 70.0d0 * (grnd() - 0.5d0)) * cdr
        else if (tmp .lt. 0.7d0) then
          vlvr(j) = (-150.0d0 + 70.0d0 * (grnd() - 0.5d0)) * cdr
        else if (tmp .lt. 0.85d0) then
          vlvr(j) = (150.0d0 + 70.0d0 * (grnd() - 0.5d0)) * cdr
        else
          vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
        end if
      end do
      call helix(nhel, mhel, nbet, mbet)
      write(*,'(a,i0)') '@vlvr double ', nvr
      do i = 1, nvr
        write(*,'(es25.17)') vlvr(idvr(i))
      end do
      write(*,'(a,i0)') '@helix int 4'
      write(*,'(i0)') nhel
      write(*,'(i0)') mhel
      write(*,'(i0)') nbet
      write(*,'(i0)') mbet
    end do

  case ('obs')
    ! observables of conformations drawn as in 'helix' (coordinates built by
    ! energy()): hbond of every chain (mhb and the type table), interhbond
    ! (mmhb), zimmer, and outpdb(0) of the first two into <prefix>_k.pdb
    call setup(2, .false.)
    call iarg(5, iseed)
    call iarg(6, nconf)
    call get_command_argument(7, wfile)
    if (command_argument_count() .ge. 9) then
      call get_command_argument(8, arg)
      call rarg(9, boxsize)
      call read_gbpr(arg)
    end if
    call sgrnd(iseed)
    write(*,'(a,i0)') '@nconf int 1'
    write(*,'(i0)') nconf
    do k = 1, nconf
      do i = 1, nvr
        j = idvr(i)
        tmp = grnd()
        if (tmp .lt. 0.35d0) then
          vlvr(j) = (-70.0d0 + 70.0d0 * (grnd() - 0.5d0)) * cdr
        else if (tmp .lt. 0.7d0) then
          vlvr(j) = (-150.0d0 + 70.0d0 * (grnd() - 0.5d0)) * cdr
        else if (tmp .lt. 0.85d0) then
          vlvr(j) = (150.0d0 + 70.0d0 * (grnd() - 0.5d0)) * cdr
        else
          vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
        end if
      end do
      e = energy()
      write(*,'(a,i0)') '@vlvr double ', nvr
      do i = 1, nvr
        write(*,'(es25.17)') vlvr(idvr(i))
      end do
      do kk = 1, ntlml
        call hbond(kk, nhel, 0)
        write(*,'(a,i0)') '@hbond int ', 2 + 2*ntyhb
        write(*,'(i0)') nhel
        write(*,'(i0)') ntyhb
        write(*,'(i0)') (ixtyhb(i), i=1,ntyhb)
        write(*,'(i0)') (nutyhb(i), i=1,ntyhb)
      end do
      call interhbond(nhel)
      write(*,'(a,i0)') '@interhb int ', 1 + ntlml*ntlml
      write(*,'(i0)') nhel
      write(*,'(i0)') ((mmhb(i,j), i=1,ntlml), j=1,ntlml)
      call zimmer(irsml2(ntlml))
      write(*,'(a,i0)') '@zimm int ', irsml2(ntlml)
      write(*,'(i0)') (ichar(zimm(i:i)), i=1,irsml2(ntlml))
      if (k .le. 2) then
        write(arg, '(a,a,i0,a)') trim(wfile), '_', k, '.pdb'
        call outpdb(0, trim(arg))
      end if
    end do

  case ('canon')
    call setup(2, .true.)
    call iarg(6, iseed)
    call rarg(7, tmp)
    call iarg(8, nsw)
    call sgrnd(iseed)
    beta = 1.0d0 / (tmp * ggasc)
    eolx = energy()
    acz = 0.0d0
    write(*,'(a,i0)') '@canon_e0 double 1'
    write(*,'(es25.17)') eolx
    do k = 1, nsw
      call metropolis(eolx, acz, can_weight)
      write(*,'(a,i0)') '@sweep double 2'
      write(*,'(es25.17)') eolx
      write(*,'(es25.17)') acz
    end do
    write(*,'(a,i0)') '@vlvr double ', nvr
    do i = 1, nvr
      write(*,'(es25.17)') vlvr(idvr(i))
    end do
    write(*,'(a,i0)') '@mti int 1'
    write(*,'(i0)') mtstate_pos()

  case ('grad')
    call setup(2, .true.)
    call iarg(6, nconf)
    call iarg(7, iseed)
    call iarg(8, iepsd)
    epsd = iepsd .ne. 0
    call sgrnd(iseed)
    nvta = ivrml1(ntlml) + nvrml(ntlml) - 1
    do k = 0, nconf
      if (k .gt. 0) then
        do i = 1, nvr
          j = idvr(i)
          vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
        end do
      end if
      write(*,'(a,i0)') '@vlvr double ', nvta
      write(*,'(es25.17)') (vlvr(i), i=1,nvta)
      call gradient()
      write(*,'(a,i0)') '@grad_energy double 5'
      write(*,'(es25.17)') eysm, eyel, eyvw, eyhb, eyvr
      write(*,'(a,i0)') '@gdeyvr double ', nvta
      write(*,'(es25.17)') (gdeyvr(i), i=1,nvta)
    end do

  case ('minim')
    call setup(2, .true.)
    call iarg(6, imin)
    call iarg(7, maxit)
    call rarg(8, epsmin)
    call iarg(9, nconf)
    call iarg(10, iseed)
    call sgrnd(iseed)
    nvta = ivrml1(ntlml) + nvrml(ntlml) - 1
    do k = 0, nconf
      if (k .gt. 0) then
        do i = 1, nvr
          j = idvr(i)
          vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
        end do
      end if
      write(*,'(a,i0)') '@vlvr double ', nvta
      write(*,'(es25.17)') (vlvr(i), i=1,nvta)
      call drv_minim(imin, maxit, epsmin, n)
      write(*,'(a,i0)') '@min_vlvr double ', nvta
      write(*,'(es25.17)') (vlvr(i), i=1,nvta)
      write(*,'(a,i0)') '@min_energy double 6'
      write(*,'(es25.17)') eysm, eyel, eyvw, eyhb, eyvr, eysl
      write(*,'(a,i0)') '@min_nfun int 1'
      write(*,'(i0)') n
    end do

  case ('rmsd')
    call setup(2, .true.)
    call iarg(6, iseed)
    call sgrnd(iseed)
    e = energy()
    nat = iatrs2(irsml2(ntlml))
    do i = 1, 3
      ang(i) = (grnd() - 0.5d0) * 2.0d0 * pi
    end do
    ! a rotation from three angles about z, y, z
    rot(1,1) = cos(ang(1))*cos(ang(2))*cos(ang(3)) - sin(ang(1))*sin(ang(3))
    rot(1,2) = -cos(ang(1))*cos(ang(2))*sin(ang(3)) - sin(ang(1))*cos(ang(3))
    rot(1,3) = cos(ang(1))*sin(ang(2))
    rot(2,1) = sin(ang(1))*cos(ang(2))*cos(ang(3)) + cos(ang(1))*sin(ang(3))
    rot(2,2) = -sin(ang(1))*cos(ang(2))*sin(ang(3)) + cos(ang(1))*cos(ang(3))
    rot(2,3) = sin(ang(1))*sin(ang(2))
    rot(3,1) = -sin(ang(2))*cos(ang(3))
    rot(3,2) = sin(ang(2))*sin(ang(3))
    rot(3,3) = cos(ang(2))
    do i = 1, nat
      px = rot(1,1)*xat(i) + rot(1,2)*yat(i) + rot(1,3)*zat(i)
      py = rot(2,1)*xat(i) + rot(2,2)*yat(i) + rot(2,3)*zat(i)
      pz = rot(3,1)*xat(i) + rot(3,2)*yat(i) + rot(3,3)*zat(i)
      xrf(i) = px + 3.0d0 + (grnd() - 0.5d0) * 0.6d0
      yrf(i) = py - 1.5d0 + (grnd() - 0.5d0) * 0.6d0
      zrf(i) = pz + 0.5d0 + (grnd() - 0.5d0) * 0.6d0
      ixat(i) = i
      if (mod(i, 11) .eq. 5) ixat(i) = 0
    end do
    write(*,'(a,i0)') '@rmsd_ref double ', 3*nat
    write(*,'(es25.17)') (xrf(i), yrf(i), zrf(i), i=1,nat)
    write(*,'(a,i0)') '@rmsd_ixat int ', nat
    write(*,'(i0)') (ixat(i), i=1,nat)
    ir1s(1) = 1
    ir2s(1) = irsml2(1) - irsml1(1) + 1
    ir1s(2) = 2
    ir2s(2) = min(4, ir2s(1))
    write(*,'(a,i0)') '@rmsd double ', 6
    do isl = 0, 2
      do q = 1, 2
        write(*,'(es25.17)') rmsdfun(1, ir1s(q), ir2s(q), ixat, xrf, yrf, zrf, isl)
      end do
    end do
    write(*,'(a,i0)') '@xyz double ', 3*nat
    write(*,'(es25.17)') (xat(i), yat(i), zat(i), i=1,nat)

  case ('metro')
    call setup(2, .true.)
    call iarg(6, iseed)
    call rarg(7, tmp)
    call iarg(8, nsw)
    call iarg(9, j)
    upchswitch = j
    call rarg(10, bgsprob)
    call iarg(11, j)
    rndord = (j .ne. 0)
    call get_command_argument(12, wfile)
    if (trim(wfile) .ne. '-') then
      open(76, file=trim(wfile), status='old')
      read(76, *) kmn, kmx, webin
      allocate(wb(kmn:kmx), wal(kmn:kmx))
      do kk = kmn, kmx
        read(76, *) wb(kk), wal(kk)
      end do
      close(76)
      wxmin = dble(kmn)
      wxmax = dble(kmx)
    end if
    call sgrnd(iseed)
    beta = 1.0d0 / (tmp * ggasc)
    eolx = energy()
    acz = 0.0d0
    ncalls = 0
    nacalls = 0
    write(*,'(a,i0)') '@canon_e0 double 1'
    write(*,'(es25.17)') eolx
    do k = 1, nsw
      if (trim(wfile) .ne. '-') then
        call metropolis(eolx, acz, drv_muca2)
      else
        call metropolis(eolx, acz, can_weight)
      end if
      write(*,'(a,i0)') '@sweep double 2'
      write(*,'(es25.17)') eolx
      write(*,'(es25.17)') acz
    end do
    write(*,'(a,i0)') '@vlvr double ', nvr
    do i = 1, nvr
      write(*,'(es25.17)') vlvr(idvr(i))
    end do
    write(*,'(a,i0)') '@mti int 1'
    write(*,'(i0)') mtstate_pos()
    write(*,'(a,i0)') '@ncalls int 10'
    write(*,'(i0)') (ncalls(i), i=1,5)
    write(*,'(i0)') (nacalls(i), i=1,5)

  case ('partem')
    call setup(2, .false.)
    call iarg(5, iseed)
    call iarg(6, nrep)
    call iarg(7, nequi)
    call iarg(8, nswp)
    call sgrnd(iseed)
    iswt = 1
    if (command_argument_count() .ge. 9) call iarg(9, iswt)
    call partem_s(nrep, nequi, nswp, 1, nswp, .true., iswt)
    write(*,'(a,i0)') '@done int 1'
    write(*,'(i0)') 1

  case ('muca')
    call setup(2, .false.)
    call iarg(5, iseed)
    call iarg(6, nsw)
    call iarg(7, nup)
    call rarg(8, tmp)
    call iarg(9, kmn)
    call iarg(10, kmx)
    call rarg(11, ebw)
    call sgrnd(iseed)
    call mulcan_par(nsw, nup, tmp, kmn, kmx, ebw, .false.)
    write(*,'(a,i0)') '@done int 1'
    write(*,'(i0)') 1

  case ('esolan')
    call get_command_argument(2, arg); libdir = arg
    call get_command_argument(3, arg); seqf = arg
    call get_command_argument(4, arg); radf = arg
    varf = ' '
    call iarg(5, iseed)
    call iarg(6, nconf)
    call setup_sol(-1)
    ! 'index radians' lines go straight into vlvr: the start energy's own
    ! setvar applies them (one rebuild from the initial torsions)
    if (trim(radf) .ne. '-') then
      open(79, file=trim(radf), status='old')
      do
        read(79, *, iostat=ios) j, vrad
        if (ios .ne. 0) exit
        vlvr(j) = vrad
      end do
      close(79)
    end if
    call sgrnd(iseed)
    e = energy()
    call dump_energy(e)
    write(*,'(a,i0)') '@mti int 1'
    write(*,'(i0)') mtstate_pos()
    call dump_coords()
    write(*,'(a,i0)') '@nconf int 1'
    write(*,'(i0)') nconf
    do k = 1, nconf
      do i = 1, nvr
        j = idvr(i)
        vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
      end do
      e = energy()
      write(*,'(a,i0)') '@vlvr double ', nvr
      do i = 1, nvr
        write(*,'(es25.17)') vlvr(idvr(i))
      end do
      call dump_energy(e)
      write(*,'(a,i0)') '@mti int 1'
      write(*,'(i0)') mtstate_pos()
    end do

  case ('multi')
    call setup(2, .false.)
    call get_command_argument(5, arg)
    call rarg(6, boxsize)
    call iarg(7, iseed)
    call iarg(8, nconf)
    call rarg(9, tmp)
    call iarg(10, nequi)
    call iarg(11, nsw)
    call read_gbpr(arg)
    call dump_params()
    call dump_topo()
    call dump_mols()
    e = energy()
    call dump_energy(e)
    call dump_inter()
    call dump_coords()
    call sgrnd(iseed)
    write(*,'(a,i0)') '@nconf int 1'
    write(*,'(i0)') nconf
    do k = 1, nconf
      do i = 1, nvr
        j = idvr(i)
        vlvr(j) = addang(pi, axvr(j)*(grnd()-0.5))
      end do
      e = energy()
      write(*,'(a,i0)') '@vlvr double ', nvr
      do i = 1, nvr
        write(*,'(es25.17)') vlvr(idvr(i))
      end do
      call dump_energy(e)
      call dump_inter()
      if (k .le. 2) call dump_coords()
    end do
    ! Metropolis from the last random conformation: nequi sweeps to leave
    ! the clash regime, then the recorded run
    beta = 1.0d0 / (tmp * ggasc)
    eolx = energy()
    acz = 0.0d0
    do k = 1, nequi
      call metropolis(eolx, acz, can_weight)
    end do
    acz = 0.0d0
    eolx = energy()
    write(*,'(a,i0)') '@gbpr_start double ', 6*ntlml
    write(*,'(es25.17)') ((gbpr(j,i), j=1,6), i=1,ntlml)
    call dump_mt()
    write(*,'(a,i0)') '@canon_e0 double 1'
    write(*,'(es25.17)') eolx
    write(*,'(a,i0)') '@vlvr_start double ', ivrml1(ntlml) + nvrml(ntlml) - 1
    write(*,'(es25.17)') (vlvr(i), i=1,ivrml1(ntlml) + nvrml(ntlml) - 1)
    write(*,'(a,i0)') '@mti_start int 1'
    write(*,'(i0)') mtstate_pos()
    do k = 1, nsw
      call metropolis(eolx, acz, can_weight)
      write(*,'(a,i0)') '@sweep double 2'
      write(*,'(es25.17)') eolx
      write(*,'(es25.17)') acz
    end do
    write(*,'(a,i0)') '@vlvr_end double ', ivrml1(ntlml) + nvrml(ntlml) - 1
    write(*,'(es25.17)') (vlvr(i), i=1,ivrml1(ntlml) + nvrml(ntlml) - 1)
    write(*,'(a,i0)') '@gbpr_end double ', 6*ntlml
    write(*,'(es25.17)') ((gbpr(j,i), j=1,6), i=1,ntlml)
    write(*,'(a,i0)') '@mti int 1'
    write(*,'(i0)') mtstate_pos()

  case default
    write(*,*) 'unknown mode ', trim(mode)
    stop 2
  end select

contains

  subroutine iarg(pos, val)
    integer, intent(in) :: pos
    integer, intent(out) :: val
    character(len=256) :: s
    call get_command_argument(pos, s)
    read(s, *) val
  end subroutine

  subroutine rarg(pos, val)
    integer, intent(in) :: pos
    double precision, intent(out) :: val
    character(len=256) :: s
    call get_command_argument(pos, s)
    read(s, *) val
  end subroutine

  ! minim.f:35-160 for ireg = 0 without its log output: minim writes its
  ! headings with formats of several records into one internal record,
  ! which stops under this compiler (minim.f:68), so the driver does its
  ! steps (gradient at the start, the free variables with unit scales, the
  ! reference's own minqsn or mincjg, which call the reference's move, then
  ! energy() and the variables wrapped into [-pi, pi]).  nop: function calls.
  subroutine drv_minim(imin, maxit, acc, nop)
    integer, intent(in) :: imin, maxit
    double precision, intent(in) :: acc
    integer, intent(out) :: nop
    double precision :: vlvrn(mxvr), gdvr(mxvr), scl(mxvr), esm, vr, energy
    double precision, allocatable :: w(:)
    integer :: i, nv, n, n1, n2, n3, n4, n5, n6, ntlvr
    allocate(w(mxvr*(mxvr+13)/2))
    call gradient()
    ntlvr = ivrml1(ntlml) + nvrml(ntlml) - 1
    nv = 0
    do i = 1, ntlvr
      if (.not. fxvr(i)) then
        nv = nv + 1
        vlvrn(nv) = vlvr(i)
        scl(nv) = 1.d0
        gdvr(nv) = gdeyvr(i)
      end if
    end do
    n = nv
    esm = eysm
    if (imin .eq. 1) then
      n1 = 1 + (n*(n+1))/2
      n2 = n1 + n
      n3 = n2 + n
      n4 = n3 + n
      n5 = n4 + n
      n6 = n5 + n
      call minqsn(n, mxvr, vlvrn, esm, gdvr, scl, acc, w, w(n1), w(n2), w(n3), w(n4), w(n5), w(n6), maxit, nop)
    else
      n1 = 1 + n
      n2 = n1 + n
      n3 = n2 + n
      n4 = n3 + n
      n5 = n4 + n
      call mincjg(n, mxvr, vlvrn, esm, gdvr, acc, w, w(n1), w(n2), w(n3), w(n4), w(n5), maxit, nop)
    end if
    eysm = energy()
    nv = 0
    do i = 1, ntlvr
      if (.not. fxvr(i)) then
        nv = nv + 1
        vr = vlvrn(nv)
        if (abs(vr) .gt. pi) vr = vr - sign(pi2, vr)
        vlvr(i) = vr
      end if
    end do
    deallocate(w)
  end subroutine

  ! ECEPP/3 set-up as universe.py does it, then init_molecule(1,'nh2','cooh').
  subroutine setup(pos, hasvar)
    integer, intent(in) :: pos
    logical, intent(in) :: hasvar
    character(len=256) :: s
    character(len=80) :: nov
    integer :: p
    p = pos
    call get_command_argument(p, s);     libdir = s
    call get_command_argument(p+1, s);   seqf = s
    if (hasvar) then
      call get_command_argument(p+2, s); varf = s
      if (trim(varf) .eq. '-') varf = ' '
      call get_command_argument(p+3, s); read(s,*) isol
    else
      varf = ' '
      call get_command_argument(p+2, s); read(s,*) isol
    end if
    flex = .false.
    sh2 = .false.
    epsd = .false.
    ientyp = 0
    tesgrd = .false.
    itysol = isol
    isolscl = .false.
    upchswitch = 0
    rndord = .false.
    bgsprob = 0.0d0
    ntlml = 0
    call init_energy(libdir)
    nov = ' '
    call init_molecule(1, grpn, grpc, seqf, nov)
    if (len_trim(varf) .gt. 0) call apply_vars(varf)
  end subroutine

  ! setup() with the library, sequence and variable files already in
  ! libdir/seqf/varf and the solvent model isol given
  subroutine setup_sol(isol_in)
    integer, intent(in) :: isol_in
    character(len=80) :: nov
    flex = .false.
    sh2 = .false.
    epsd = .false.
    ientyp = 0
    tesgrd = .false.
    itysol = isol_in
    isolscl = .false.
    upchswitch = 0
    rndord = .false.
    bgsprob = 0.0d0
    ntlml = 0
    call init_energy(libdir)
    nov = ' '
    call init_molecule(1, grpn, grpc, seqf, nov)
    if (len_trim(varf) .gt. 0) call apply_vars(varf)
  end subroutine

  ! Applies 'index value_deg fixed' lines (variable index over all variables,
  ! angle in degrees, fixed flag 0/1), then rebuilds and recounts the free
  ! variables the way init_molecule.f:112-131 does after redvar.
  subroutine apply_vars(fname)
    character(len=*), intent(in) :: fname
    integer :: iv, ifx, ios, a, nvt
    double precision :: vdeg
    open(77, file=trim(fname), status='old')
    do
      read(77, *, iostat=ios) iv, vdeg, ifx
      if (ios .ne. 0) exit
      vlvr(iv) = vdeg * cdr
      fxvr(iv) = (ifx .ne. 0)
    end do
    close(77)
    call setvar(1, vlvr)
    nvt = ivrml1(ntlml) + nvrml(ntlml) - 1
    nvr = 0
    do a = 1, nvt
      if (.not. fxvr(a)) then
        nvr = nvr + 1
        idvr(nvr) = a
      end if
      vlvr(a) = toat(iatvr(a))
      olvlvr(a) = vlvr(a)
      if (fxvr(a)) axvr(a) = 0.0d0
    end do
  end subroutine

  subroutine read_gbpr(fname)
    character(len=*), intent(in) :: fname
    integer :: iml, kk, ios
    double precision :: v
    open(78, file=trim(fname), status='old')
    do
      read(78, *, iostat=ios) iml, kk, v
      if (ios .ne. 0) exit
      gbpr(kk, iml) = v
    end do
    close(78)
  end subroutine

  subroutine dump_mols()
    integer :: m
    write(*,'(a,i0)') '@irsml1 int ', ntlml
    write(*,'(i0)') (irsml1(m), m=1,ntlml)
    write(*,'(a,i0)') '@irsml2 int ', ntlml
    write(*,'(i0)') (irsml2(m), m=1,ntlml)
    write(*,'(a,i0)') '@ivrml1 int ', ntlml
    write(*,'(i0)') (ivrml1(m), m=1,ntlml)
    write(*,'(a,i0)') '@nvrml int ', ntlml
    write(*,'(i0)') (nvrml(m), m=1,ntlml)
    write(*,'(a,i0)') '@gbpr double ', 6*ntlml
    write(*,'(es25.17)') ((gbpr(j,m), j=1,6), m=1,ntlml)
  end subroutine

  subroutine dump_inter()
    write(*,'(a,i0)') '@inter double 4'
    write(*,'(es25.17)') eysmi, eyeli, eyvwi, eyhbi
  end subroutine

  subroutine dump_mt()
    integer :: mtl(0:623), mtil, q
    common /myblock/ mtil, mtl
    write(*,'(a,i0)') '@mt_state int 625'
    write(*,'(i0)') (mtl(q), q=0,623)
    write(*,'(i0)') mtil
  end subroutine

  integer function mtstate_pos()
    integer :: mtl(0:623), mtil
    common /myblock/ mtil, mtl
    mtstate_pos = mtil
  end function

  subroutine dump_params()
    integer :: a, b
    write(*,'(a,i0)') '@conv double 1'
    write(*,'(es25.17)') conv
    write(*,'(a,i0)') '@cij double ', mxtyat*mxtyat
    write(*,'(es25.17)') ((cij(a,b), b=1,mxtyat), a=1,mxtyat)
    write(*,'(a,i0)') '@aij double ', mxtyat*mxtyat
    write(*,'(es25.17)') ((aij(a,b), b=1,mxtyat), a=1,mxtyat)
    write(*,'(a,i0)') '@a14 double ', mxtyat*mxtyat
    write(*,'(es25.17)') ((a14(a,b), b=1,mxtyat), a=1,mxtyat)
    write(*,'(a,i0)') '@chb double ', mxtyat*mxtyat
    write(*,'(es25.17)') ((chb(a,b), b=1,mxtyat), a=1,mxtyat)
    write(*,'(a,i0)') '@ahb double ', mxtyat*mxtyat
    write(*,'(es25.17)') ((ahb(a,b), b=1,mxtyat), a=1,mxtyat)
    write(*,'(a,i0)') '@ihbty int ', mxtyat*mxtyat
    write(*,'(i0)') ((ihbty(a,b), b=1,mxtyat), a=1,mxtyat)
    write(*,'(a,i0)') '@e0to double ', mxtyto
    write(*,'(es25.17)') (e0to(a), a=1,mxtyto)
    write(*,'(a,i0)') '@sgto double ', mxtyto
    write(*,'(es25.17)') (sgto(a), a=1,mxtyto)
    write(*,'(a,i0)') '@rnto double ', mxtyto
    write(*,'(es25.17)') (rnto(a), a=1,mxtyto)
    write(*,'(a,i0)') '@npnt int 1'
    write(*,'(i0)') npnt
    if (npnt .gt. 0) then
      write(*,'(a,i0)') '@spoint double ', 3*npnt
      write(*,'(es25.17)') ((spoint(a,b), b=1,3), a=1,npnt)
    end if
  end subroutine

  subroutine dump_topo()
    integer :: nat, nvt, nms, a, b
    nat = iatrs2(irsml2(ntlml))
    nvt = ivrml1(ntlml) + nvrml(ntlml) - 1
    nms = imsml1(ntlml) + nmsml(ntlml) - 1
    write(*,'(a,i0)') '@ntlml int 1'
    write(*,'(i0)') ntlml
    write(*,'(a,i0)') '@nres int 1'
    write(*,'(i0)') irsml2(ntlml)
    write(*,'(a,i0)') '@seq str ', irsml2(ntlml)
    write(*,'(a)') (seq(a), a=1,irsml2(ntlml))
    write(*,'(a,i0)') '@iatrs1 int ', irsml2(ntlml)
    write(*,'(i0)') (iatrs1(a), a=1,irsml2(ntlml))
    write(*,'(a,i0)') '@iatrs2 int ', irsml2(ntlml)
    write(*,'(i0)') (iatrs2(a), a=1,irsml2(ntlml))
    write(*,'(a,i0)') '@ixatrs int ', irsml2(ntlml)
    write(*,'(i0)') (ixatrs(a), a=1,irsml2(ntlml))
    write(*,'(a,i0)') '@ivrrs1 int ', irsml2(ntlml)
    write(*,'(i0)') (ivrrs1(a), a=1,irsml2(ntlml))
    write(*,'(a,i0)') '@nvrrs int ', irsml2(ntlml)
    write(*,'(i0)') (nvrrs(a), a=1,irsml2(ntlml))
    write(*,'(a,i0)') '@nat int 1'
    write(*,'(i0)') nat
    write(*,'(a,i0)') '@nmat str ', nat
    write(*,'(a)') (nmat(a), a=1,nat)
    write(*,'(a,i0)') '@ityat int ', nat
    write(*,'(i0)') (ityat(a), a=1,nat)
    write(*,'(a,i0)') '@iowat int ', nat
    write(*,'(i0)') (iowat(a), a=1,nat)
    write(*,'(a,i0)') '@iyowat int ', nat
    write(*,'(i0)') (iyowat(a), a=1,nat)
    write(*,'(a,i0)') '@nbdat int ', nat
    write(*,'(i0)') (nbdat(a), a=1,nat)
    write(*,'(a,i0)') '@ibdat int ', 3*nat
    write(*,'(i0)') ((ibdat(b,a), b=1,3), a=1,nat)
    write(*,'(a,i0)') '@iybdat int ', 3*nat
    write(*,'(i0)') ((iybdat(b,a), b=1,3), a=1,nat)
    write(*,'(a,i0)') '@cgat double ', nat
    write(*,'(es25.17)') (cgat(a), a=1,nat)
    write(*,'(a,i0)') '@blat double ', nat
    write(*,'(es25.17)') (blat(a), a=1,nat)
    write(*,'(a,i0)') '@baat double ', nat
    write(*,'(es25.17)') (baat(a), a=1,nat)
    write(*,'(a,i0)') '@toat double ', nat
    write(*,'(es25.17)') (toat(a), a=1,nat)
    write(*,'(a,i0)') '@ixmsat int ', nat
    write(*,'(i0)') (ixmsat(a), a=1,nat)
    write(*,'(a,i0)') '@sigma double ', nat
    write(*,'(es25.17)') (sigma(a), a=1,nat)
    write(*,'(a,i0)') '@rvdw double ', nat
    write(*,'(es25.17)') (rvdw(a), a=1,nat)
    write(*,'(a,i0)') '@nvt int 1'
    write(*,'(i0)') nvt
    write(*,'(a,i0)') '@nmvr str ', nvt
    write(*,'(a)') (nmvr(a), a=1,nvt)
    write(*,'(a,i0)') '@iatvr int ', nvt
    write(*,'(i0)') (iatvr(a), a=1,nvt)
    write(*,'(a,i0)') '@ityvr int ', nvt
    write(*,'(i0)') (ityvr(a), a=1,nvt)
    write(*,'(a,i0)') '@iclvr int ', nvt
    write(*,'(i0)') (iclvr(a), a=1,nvt)
    write(*,'(a,i0)') '@iorvr int ', nvt
    write(*,'(i0)') (iorvr(a), a=1,nvt)
    write(*,'(a,i0)') '@imsvr1 int ', nvt
    write(*,'(i0)') (imsvr1(a), a=1,nvt)
    write(*,'(a,i0)') '@imsvr2 int ', nvt
    write(*,'(i0)') (imsvr2(a), a=1,nvt)
    write(*,'(a,i0)') '@iadvr1 int ', nvt
    write(*,'(i0)') (iadvr1(a), a=1,nvt)
    write(*,'(a,i0)') '@iadvr2 int ', nvt
    write(*,'(i0)') (iadvr2(a), a=1,nvt)
    write(*,'(a,i0)') '@nadml int 1'
    write(*,'(i0)') nadml(ntlml)
    write(*,'(a,i0)') '@ladvr int ', nadml(ntlml)
    write(*,'(i0)') (ladvr(a), a=1,nadml(ntlml))
    write(*,'(a,i0)') '@fxvr int ', nvt
    write(*,'(i0)') (merge(1,0,fxvr(a)), a=1,nvt)
    write(*,'(a,i0)') '@axvr double ', nvt
    write(*,'(es25.17)') (axvr(a), a=1,nvt)
    write(*,'(a,i0)') '@vlvr_all double ', nvt
    write(*,'(es25.17)') (vlvr(a), a=1,nvt)
    write(*,'(a,i0)') '@nvr int 1'
    write(*,'(i0)') nvr
    write(*,'(a,i0)') '@idvr int ', nvr
    write(*,'(i0)') (idvr(a), a=1,nvr)
    write(*,'(a,i0)') '@nms int 1'
    write(*,'(i0)') nms
    write(*,'(a,i0)') '@latms1 int ', nms
    write(*,'(i0)') (latms1(a), a=1,nms)
    write(*,'(a,i0)') '@latms2 int ', nms
    write(*,'(i0)') (latms2(a), a=1,nms)
    write(*,'(a,i0)') '@ivwat1 int ', nat
    write(*,'(i0)') (ivwat1(a), a=1,nat)
    write(*,'(a,i0)') '@ivwat2 int ', nat
    write(*,'(i0)') (ivwat2(a), a=1,nat)
    write(*,'(a,i0)') '@nvw int 1'
    write(*,'(i0)') nvwml(ntlml) + ivwml1(ntlml) - 1
    write(*,'(a,i0)') '@lvwat1 int ', nvwml(ntlml) + ivwml1(ntlml) - 1
    write(*,'(i0)') (lvwat1(a), a=1,nvwml(ntlml) + ivwml1(ntlml) - 1)
    write(*,'(a,i0)') '@lvwat2 int ', nvwml(ntlml) + ivwml1(ntlml) - 1
    write(*,'(i0)') (lvwat2(a), a=1,nvwml(ntlml) + ivwml1(ntlml) - 1)
    write(*,'(a,i0)') '@i14at1 int ', nat
    write(*,'(i0)') (i14at1(a), a=1,nat)
    write(*,'(a,i0)') '@i14at2 int ', nat
    write(*,'(i0)') (i14at2(a), a=1,nat)
    write(*,'(a,i0)') '@n14 int 1'
    write(*,'(i0)') n14ml(ntlml) + i14ml1(ntlml) - 1
    write(*,'(a,i0)') '@l14at int ', n14ml(ntlml) + i14ml1(ntlml) - 1
    write(*,'(i0)') (l14at(a), a=1,n14ml(ntlml) + i14ml1(ntlml) - 1)
  end subroutine

  subroutine dump_energy(etot)
    double precision, intent(in) :: etot
    write(*,'(a,i0)') '@energy double 7'
    write(*,'(es25.17)') etot, eysm, eyel, eyvw, eyhb, eyvr, eysl
    write(*,'(a,i0)') '@asa double 2'
    write(*,'(es25.17)') asa, vdvol
  end subroutine

  subroutine dump_coords()
    integer :: nat, a
    nat = iatrs2(irsml2(ntlml))
    write(*,'(a,i0)') '@xyz double ', 3*nat
    write(*,'(es25.17)') (xat(a), yat(a), zat(a), a=1,nat)
  end subroutine

end program ref_driver
