! host_io.f90 -- readers of the LS-EVPFFT input files and the CSV writers.
!
! Readers restate load_input / data_crystal / data_crystal_elast / data_grain
! (lsevpfft/src/IO_functions.f90:6-365, 367-637, 640-719, 722-812) with
! Fortran list-directed reads, so the same fft.in / cupl2.sx / cuel.sx /
! microstructure files are accepted.  Limits are dynamic (no nsysmx / ngrmx).
! Writers restate write_results (IO_functions.f90:1237-1303) and
! write_diagnostics (types.f90:760-834).
module host_io
    use iso_c_binding
    use creep_hip_iface
    implicit none

    type phase_t
        logical :: igas = .false.
        integer :: nsyst = 0, nmodes = 0
        integer, allocatable :: nsm(:)
        real(8), allocatable :: dnca(:, :), dbca(:, :)   ! (3, ns)
        real(8), allocatable :: tau(:, :)                ! (ns, 3)
        real(8), allocatable :: thet(:, :)               ! (ns, 0:1)
        real(8), allocatable :: nrs(:), gamd0(:), hpfac(:)
        real(8), allocatable :: hard(:, :)
        real(8) :: c(3, 3, 3, 3) = 0.0d0
    end type phase_t

    type input_t
        integer :: nph = 0, npts1 = 0, npts2 = 0, npts3 = 0, nproc = 0, idispini = 0
        real(8) :: complgas = 0.0d0
        type(phase_t), allocatable :: phase(:)
        real(8), allocatable :: euler(:, :)     ! (3, V) degrees
        integer(c_int), allocatable :: grain(:), ph(:)
        type(cg_process), allocatable :: proc(:)
        ! initial displacement field (init_defgrad, IO_functions.f90:850-880): 3 values per cell
        ! (idispini 1) or per node (2), x fastest, as the file lists them
        real(c_double), allocatable :: dispini(:)
    end type input_t

    ! eigenstrain rates of ithermo = 1 (IO_functions.f90:316-341): one 3x3 tensor per grain from
    ! the file the last such process names (every process assigns the grain table to the voxels,
    ! and the table holds what was read last)
    real(8), allocatable :: eth_grain(:, :, :)   ! (3, 3, ngrains)
    logical :: have_thermo = .false.

contains

    subroutine fatal(msg)
        character(len=*), intent(in) :: msg
        write (*, '(A)') trim(msg)
        stop 1
    end subroutine fatal

    ! A failed library call: status 1 for a singular system (the reference's `stop 1` in
    ! lu_inverse / lu_eqsystem, tensor_functions.f90:323-334, 353-356) and for runtime
    ! errors, status 2 for non-finite fixed-point errors.
    subroutine fatal_rc(msg, rc)
        character(len=*), intent(in) :: msg
        integer, intent(in) :: rc
        write (*, '(A)') trim(msg)
        if (rc == 2) stop 2
        stop 1
    end subroutine fatal_rc

    subroutine load_input(fname, inp)
        character(len=*), intent(in) :: fname
        type(input_t), intent(out) :: inp
        character(len=256) :: prosa, filetext, fpl, fel, filedispl
        integer :: iph, ii, i, imicrosave, ipr
        real(8) :: delt(3)

        open (unit=11, file=fname, status='old')
        read (11, *) inp%nph
        read (11, *) inp%npts1, inp%npts2, inp%npts3
        read (11, *) delt(1), delt(2), delt(3)
        read (11, '(a)') prosa
        read (11, '(a)') filetext
        allocate (inp%phase(inp%nph))
        do iph = 1, inp%nph
            read (11, '(a)') prosa
            read (11, *) ii
            inp%phase(iph)%igas = (ii /= 0)
            read (11, '(a)') prosa
            read (11, '(a)') fpl
            read (11, '(a)') fel
            if (.not. inp%phase(iph)%igas) then
                open (unit=12, file=trim(fpl), status='old')
                call read_crystal(inp%phase(iph))
                close (12)
                open (unit=12, file=trim(fel), status='old')
                call read_elastic(inp%phase(iph))
                close (12)
            end if
        end do
        open (unit=12, file=trim(filetext), status='old')
        call read_grains(inp)
        close (12)
        read (11, '(a)') prosa
        read (11, *) i, imicrosave
        read (11, *) inp%complgas
        read (11, '(a)') prosa
        read (11, *) inp%idispini
        if (inp%idispini > 0) then
            read (11, '(A)') filedispl
            call read_displacement(trim(adjustl(filedispl)), inp)
        end if
        read (11, *) inp%nproc
        if (inp%nproc < 1) call fatal('cannot have less than one processes')
        allocate (inp%proc(inp%nproc))
        do ipr = 1, inp%nproc
            call read_process(inp%proc(ipr))
        end do
        close (11)
    end subroutine load_input

    ! data_crystal (IO_functions.f90:367-637), cubic/orthorhombic and hexagonal/trigonal
    subroutine read_crystal(p)
        type(phase_t), intent(inout) :: p
        character(len=256) :: prosa
        character(len=3) :: icryst
        real(8) :: cdim(3), covera, nrsx, gamd0x, twshx, tau0f, tau0b, tau1, th0, th1, hpf
        real(8) :: sn(3), sb(3), dn(3), db(3)
        real(8), allocatable :: hself(:), hlat(:, :)
        real(8), allocatable :: t_dn(:, :), t_db(:, :), t_tau(:, :), t_th(:, :), t_nrs(:), t_g0(:), t_hp(:)
        integer, allocatable :: mode(:), t_mode(:)
        integer :: nmodesx, nmodes, nm, modex, nsmx, isectw, kount, j, k, m, ns, i, isn(4), isb(4)
        logical :: hexa

        read (12, '(a)') prosa
        read (12, '(a)') icryst
        read (12, *) (cdim(i), i=1, 3)
        covera = cdim(3) / cdim(1)
        read (12, *) nmodesx
        read (12, *) nmodes
        allocate (mode(nmodes), hself(nmodesx), hlat(nmodesx, nmodesx))
        read (12, *) (mode(i), i=1, nmodes)
        allocate (t_dn(3, 1024), t_db(3, 1024), t_tau(1024, 3), t_th(1024, 2), t_nrs(1024), t_g0(1024), &
                  t_hp(1024), t_mode(1024))
        hexa = (icryst == 'HEX' .or. icryst == 'TRI' .or. icryst == 'hex' .or. icryst == 'tri')
        if (.not. hexa .and. .not. (icryst == 'CUB' .or. icryst == 'ORT' .or. icryst == 'cub' .or. icryst == 'ort')) &
            call fatal(' cannot identify the crystal symmetry of phase')
        ns = 0
        kount = 1
        do nm = 1, nmodesx
            read (12, '(a)') prosa
            read (12, *) modex, nsmx, nrsx, gamd0x, twshx, isectw
            read (12, *) tau0f, tau0b, tau1, th0, th1, hpf
            read (12, *) hself(nm), (hlat(nm, j), j=1, nmodesx)
            if (kount > nmodes) then
                do j = 1, nsmx
                    read (12, *)
                end do
                cycle
            end if
            if (modex /= mode(kount)) then
                do j = 1, nsmx
                    read (12, *)
                end do
                cycle
            end if
            if (th0 < th1) call fatal(' initial hardening lower than final hardening')
            if (tau1 <= 1.d-6) then
                tau1 = 1.d-6
                th0 = th1
            end if
            if (twshx > 1.0d-15) call fatal(' twinning modes are not supported (ntwmmx=0)')
            do j = 1, nsmx
                if (hexa) then
                    read (12, *) (isn(k), k=1, 4), (isb(k), k=1, 4)
                    sn(1) = dble(isn(1))
                    sn(2) = (dble(isn(1)) + 2.d0 * dble(isn(2))) / sqrt(3.d0)
                    sn(3) = dble(isn(4)) / covera
                    sb(1) = 1.5d0 * dble(isb(1))
                    sb(2) = (dble(isb(1)) / 2.d0 + dble(isb(2))) * sqrt(3.d0)
                    sb(3) = dble(isb(4)) * covera
                else
                    read (12, *) (isn(k), k=1, 3), (isb(k), k=1, 3)
                    do m = 1, 3
                        sn(m) = dble(isn(m)) / cdim(m)
                        sb(m) = dble(isb(m)) * cdim(m)
                    end do
                end if
                dn = sn / sqrt(sum(sn * sn))
                db = sb / sqrt(sum(sb * sb))
                where (abs(dn) < 1.d-3) dn = 0.d0
                where (abs(db) < 1.d-3) db = 0.d0
                if (sum(dn * db) >= 1.d-3) call fatal(' slip system is not orthogonal')
                ns = ns + 1
                if (ns > size(t_nrs)) call fatal(' too many slip systems')
                t_dn(:, ns) = dn
                t_db(:, ns) = db
                t_tau(ns, :) = [tau0f, tau0b, tau1]
                t_th(ns, :) = [th0, th1]
                t_nrs(ns) = nrsx
                t_g0(ns) = gamd0x
                t_hp(ns) = hpf
                t_mode(ns) = kount
            end do
            kount = kount + 1
        end do
        p%nsyst = ns
        p%nmodes = nmodes
        allocate (p%nsm(nmodes), p%dnca(3, ns), p%dbca(3, ns), p%tau(ns, 3), p%thet(ns, 0:1), p%nrs(ns), &
                  p%gamd0(ns), p%hpfac(ns), p%hard(ns, ns))
        do m = 1, nmodes
            p%nsm(m) = count(t_mode(1:ns) == m)
        end do
        p%dnca = t_dn(:, 1:ns)
        p%dbca = t_db(:, 1:ns)
        p%tau = t_tau(1:ns, :)
        p%thet(:, 0) = t_th(1:ns, 1)
        p%thet(:, 1) = t_th(1:ns, 2)
        p%nrs = t_nrs(1:ns)
        p%gamd0 = t_g0(1:ns)
        p%hpfac = t_hp(1:ns)
        ! hard(i,j) = hlatex(mode_i, mode_j), hselfx(mode_i) on the diagonal, modes taken in the
        ! order of the active list (IO_functions.f90:468-471, 578-591)
        do i = 1, ns
            do j = 1, ns
                p%hard(i, j) = hlat(mode(t_mode(i)), mode(t_mode(j)))
            end do
            p%hard(i, i) = hself(mode(t_mode(i)))
        end do
    end subroutine read_crystal

    ! data_crystal_elast (IO_functions.f90:640-719)
    subroutine read_elastic(p)
        type(phase_t), intent(inout) :: p
        integer :: iso, i, j, k, l, a, b
        real(8) :: c66(6, 6), young, tnu, tmu, tla, d(3, 3)
        integer, parameter :: ijv(6, 2) = reshape([1, 2, 3, 2, 1, 1, 1, 2, 3, 3, 3, 2], [6, 2])
        read (12, *) iso
        p%c = 0.d0
        if (iso == 0) then
            do i = 1, 6
                read (12, *) (c66(i, j), j=1, 6)
            end do
            do a = 1, 6
                do b = 1, 6
                    p%c(ijv(a, 1), ijv(a, 2), ijv(b, 1), ijv(b, 2)) = c66(a, b)
                    p%c(ijv(a, 2), ijv(a, 1), ijv(b, 1), ijv(b, 2)) = c66(a, b)
                    p%c(ijv(a, 1), ijv(a, 2), ijv(b, 2), ijv(b, 1)) = c66(a, b)
                    p%c(ijv(a, 2), ijv(a, 1), ijv(b, 2), ijv(b, 1)) = c66(a, b)
                end do
            end do
        else
            read (12, *) young, tnu
            tmu = young / (2.d0 * (1.d0 + tnu))
            tla = 2.d0 * tmu * tnu / (1.d0 - 2.d0 * tnu)
            d = 0.d0
            do i = 1, 3
                d(i, i) = 1.d0
            end do
            do i = 1, 3
                do j = 1, 3
                    do k = 1, 3
                        do l = 1, 3
                            p%c(i, j, k, l) = tla * d(i, j) * d(k, l) + tmu * (d(i, k) * d(j, l) + d(i, l) * d(j, k))
                        end do
                    end do
                end do
            end do
        end if
    end subroutine read_elastic

    ! data_grain (IO_functions.f90:759-809): explicit indices, any order
    subroutine read_grains(inp)
        type(input_t), intent(inout) :: inp
        integer :: kkk, ii, jj, kk, jgr, jph
        integer(8) :: v, nv
        real(8) :: ph, th, om
        nv = int(inp%npts1, 8) * inp%npts2 * inp%npts3
        allocate (inp%euler(3, nv), inp%grain(nv), inp%ph(nv))
        do kkk = 1, int(nv)
            read (12, *) ph, th, om, ii, jj, kk, jgr, jph
            if (ii < 1 .or. ii > inp%npts1 .or. jj < 1 .or. jj > inp%npts2 .or. kk < 1 .or. kk > inp%npts3) &
                call fatal(' voxel index out of range in microstructure file')
            if (jph < 1 .or. jph > inp%nph) call fatal(' phase id out of range in microstructure file')
            v = ii + int(inp%npts1, 8) * ((jj - 1) + int(inp%npts2, 8) * (kk - 1))
            inp%euler(:, v) = [ph, th, om]
            inp%grain(v) = jgr
            inp%ph(v) = jph
        end do
    end subroutine read_grains

    ! one process block (IO_functions.f90:202-345)
    subroutine read_process(p)
        type(cg_process), intent(inout) :: p
        integer :: i, j, iud(3, 3), isc(6), irecover, isave, iwtex
        real(8) :: ud(3, 3), sc(3, 3), ds(3, 3), dvm, h
        read (11, *)
        read (11, *)
        do i = 1, 3
            read (11, *) (iud(i, j), j=1, 3)
        end do
        if (iud(1, 1) + iud(2, 2) + iud(3, 3) == 2) call fatal('check diagonal boundary conditions iudot')
        do i = 1, 3
            do j = 1, 3
                if (i /= j .and. iud(i, j) + iud(j, i) == 0) call fatal('check off-diagonal boundary conditions iudot')
            end do
        end do
        read (11, *)
        do i = 1, 3
            read (11, *) (ud(i, j), j=1, 3)
        end do
        ds = 0.5d0 * (ud + transpose(ud))
        h = (ds(1, 1) + ds(2, 2) + ds(3, 3)) / 3.d0
        dvm = sqrt(2.d0 / 3.d0 * ((ds(1, 1) - h)**2 + (ds(2, 2) - h)**2 + (ds(3, 3) - h)**2 + &
                                 2.d0 * (ds(1, 2)**2 + ds(1, 3)**2 + ds(2, 3)**2)))
        p%idsim(1) = iud(1, 1); p%idsim(2) = iud(2, 2); p%idsim(3) = iud(3, 3)
        p%idsim(4) = 0; if (iud(2, 3) == 1 .and. iud(3, 2) == 1) p%idsim(4) = 1
        p%idsim(5) = 0; if (iud(1, 3) == 1 .and. iud(3, 1) == 1) p%idsim(5) = 1
        p%idsim(6) = 0; if (iud(1, 2) == 1 .and. iud(2, 1) == 1) p%idsim(6) = 1
        read (11, *)
        read (11, *) isc(1), isc(6), isc(5)
        read (11, *) isc(2), isc(4)
        read (11, *) isc(3)
        do i = 1, 6
            if (isc(i) * p%idsim(i) /= 0 .or. isc(i) + p%idsim(i) /= 1) &
                call fatal(' check boundary condits on strain-rate and stress')
        end do
        p%iscau = isc
        read (11, *)
        sc = 0.d0
        read (11, *) sc(1, 1), sc(1, 2), sc(1, 3)
        read (11, *) sc(2, 2), sc(2, 3)
        read (11, *) sc(3, 3)
        sc(3, 2) = sc(2, 3); sc(3, 1) = sc(1, 3); sc(2, 1) = sc(1, 2)
        do i = 1, 3
            do j = 1, 3
                p%iudot(3 * (i - 1) + j) = iud(i, j)
                p%udot(3 * (i - 1) + j) = ud(i, j)
                p%scauchy(3 * (i - 1) + j) = sc(i, j)
                p%dsim(3 * (i - 1) + j) = ds(i, j)
            end do
        end do
        read (11, *)
        read (11, *) p%tdot, p%tdotref, p%tdotmin, p%devmmx
        read (11, *) p%ictrl
        p%devmapp = 0.d0
        if (p%ictrl == 0) then
            p%devmapp = p%tdot
            p%tdot = p%tdot / dvm
            p%tdotref = p%tdot
        else if (p%ictrl > 0) then
            call fatal('ictrl>0 not implemented')
        end if
        read (11, *)
        read (11, *) p%nsteps
        read (11, *) p%error, p%erroral
        read (11, *) p%itmax
        read (11, *) irecover
        read (11, *) isave
        read (11, *) p%iupdate
        read (11, *) p%iuphard
        read (11, *) iwtex
        read (11, *) p%iwfields, p%iwstep
        read (11, *) p%xc0, p%igamma
        if (.not. (p%nsteps > 0)) call fatal('nsteps must be positive')
        if (.not. (p%error > 0 .or. p%erroral > 0)) call fatal('error and erroral tolerances must be positive')
        if (.not. (p%itmax > 0)) call fatal('itmax must be positive')
        if (.not. (p%iupdate == 0 .or. p%iupdate == 1)) call fatal('iupdate must be 0 or 1')
        if (.not. (p%iuphard == 0 .or. p%iuphard == 1)) call fatal('iuphard must be 0 or 1')
        if (.not. (p%iwfields == 0 .or. p%iwfields == 1)) call fatal('iwfields must be 0 or 1')
        if (.not. (p%igamma == 0 .or. p%igamma == 1 .or. p%igamma == 2)) &
            call fatal('igamma must be 0 (continuous), 1 (discrete), or 2 (modified)')
        read (11, *) p%ithermo
        if (p%ithermo == 1) call read_thermo()
        if (irecover /= 0 .and. irecover /= 1) call fatal('irecover must be 0 or 1')
        if (isave /= 0 .and. isave /= 1) call fatal('isave must be 0 or 1')
        if (iwtex /= 0 .and. iwtex /= 1) call fatal('iwtex must be 0 or 1')
    end subroutine read_process

    ! the eigenstrain file: ngrains, then eth11 eth22 eth33 eth23 eth31 eth12 per grain
    ! the initial displacement file: one line "i j k u1 u2 u3" per cell (idispini 1) or node
    ! (idispini 2; one z layer of nodes for a 2D grid), x fastest (IO_functions.f90:850-880)
    subroutine read_displacement(fname, inp)
        character(len=*), intent(in) :: fname
        type(input_t), intent(inout) :: inp
        integer :: nrow, r, ip1, ip2, ip3, ios, n3node
        if (inp%idispini /= 1 .and. inp%idispini /= 2) call fatal('IDISPINI must be 0, 1 or 2')
        n3node = merge(1, inp%npts3 + 1, inp%npts3 == 1)
        if (inp%idispini == 1) then
            nrow = inp%npts1 * inp%npts2 * inp%npts3
        else
            nrow = (inp%npts1 + 1) * (inp%npts2 + 1) * n3node
        end if
        allocate (inp%dispini(3 * nrow))
        open (unit=13, file=fname, status='old', iostat=ios)
        if (ios /= 0) call fatal('cannot open the initial displacement file '//fname)
        do r = 1, nrow
            read (13, *, iostat=ios) ip1, ip2, ip3, inp%dispini(3 * r - 2), inp%dispini(3 * r - 1), inp%dispini(3 * r)
            if (ios /= 0) call fatal('initial displacement file '//fname//': too few lines')
        end do
        close (13)
    end subroutine read_displacement

    subroutine read_thermo()
        character(len=256) :: filethermo
        integer :: ngr, ig
        real(8) :: e(6)
        read (11, '(a)') filethermo
        open (unit=49, file=trim(adjustl(filethermo)), status='old')
        read (49, *) ngr
        if (allocated(eth_grain)) deallocate (eth_grain)
        allocate (eth_grain(3, 3, ngr))
        do ig = 1, ngr
            read (49, *) e
            eth_grain(1, 1, ig) = e(1); eth_grain(2, 2, ig) = e(2); eth_grain(3, 3, ig) = e(3)
            eth_grain(2, 3, ig) = e(4); eth_grain(3, 2, ig) = e(4)
            eth_grain(3, 1, ig) = e(5); eth_grain(1, 3, ig) = e(5)
            eth_grain(1, 2, ig) = e(6); eth_grain(2, 1, ig) = e(6)
        end do
        close (49)
        have_thermo = .true.
    end subroutine read_thermo

    ! results.csv (IO_functions.f90:1237-1303)
    subroutine write_results(dir, init, iter, imacro, ipr, sim_time, wall_time, tdot, tdot_ref, o)
        character(len=*), intent(in) :: dir
        logical, intent(in) :: init
        integer, intent(in) :: iter, imacro, ipr
        real(8), intent(in) :: sim_time, wall_time, tdot, tdot_ref
        type(cg_incr_out), intent(in) :: o
        character(len=512) :: path
        integer :: u
        path = csv_path(dir, 'results.csv')
        if (init) then
            open (newunit=u, file=trim(path), status='replace', action='write')
            write (u, '(A)') 'iter,macrostep,ipr,sim_time,wall_time,tdot,tdot_ref,'// &
                'eav11,eav22,eav33,eav23,eav13,eav12,'// &
                'sav11,sav22,sav33,sav23,sav13,sav12,'// &
                'epav11,epav22,epav33,epav23,epav13,epav12,'// &
                'eelav11,eelav22,eelav33,eelav23,eelav13,eelav12,'// &
                'vgrad11,vgrad22,vgrad33,vgrad23,vgrad13,vgrad12,'// &
                'edotp11,edotp22,edotp33,edotp23,edotp13,edotp12,'// &
                'evm,evmp,dvm,dvmp,svm,gamma,g,geff'
            close (u)
        end if
        open (newunit=u, file=trim(path), status='old', position='append', action='write')
        write (u, '(*(g0,:,","))') iter, imacro, ipr, sim_time, wall_time, tdot, tdot_ref, &
            v6(o%eavg), v6(o%sgsolavg), v6(o%epavg), v6(o%eelavg), v6(o%velgradavg), v6(o%edotpavg), &
            o%evmavg, o%evmpavg, o%dvmavg, o%dvmpavg, o%svmavg, o%gammaavg, o%gavg, o%geffavg
        flush (u)
        close (u)
    end subroutine write_results

    ! diagnostics.csv (types.f90:760-834)
    ! vals: vm_max_inc, tdotref, erre, errs, errsbc, s33_target, s33, s33_sol, detF (4), volinc (4),
    ! stretchinc (4), rotang (2), dL_step (6), dvelgradavg_mag, cond (3); imin: sol(3), gas(3)
    subroutine write_diag(dir, init, iter, imacro, ipr, sim_time, wall_time, tdot, vals, imin)
        character(len=*), intent(in) :: dir
        logical, intent(in) :: init
        integer, intent(in) :: iter, imacro, ipr, imin(6)
        real(8), intent(in) :: sim_time, wall_time, tdot, vals(32)
        character(len=512) :: path
        integer :: u
        path = csv_path(dir, 'diagnostics.csv')
        if (init) then
            open (newunit=u, file=trim(path), status='replace', action='write')
            write (u, '(A)') 'iter,macrostep,ipr,sim_time,wall_time,tdot,'// &
                'vm_max_inc,'//'tdotref,'//'erre,errs,errsbc,'//'s33_target,s33,s33_sol,'// &
                'detF_min_sol,detF_min_gas,detF_min_sol_rel,detF_min_gas_rel,'// &
                'volinc_max_sol,volinc_max_gas,volinc_max_sol_rel,volinc_max_gas_rel,'// &
                'stretchinc_max_sol,stretchinc_max_gas,stretchinc_max_sol_rel,stretchinc_max_gas_rel,'// &
                'rotang_max_sol,rotang_max_gas,'// &
                'dL_step_rms_sol,dL_step_rms_gas,dL_step_rms,dL_step_max_sol,dL_step_max_gas,dL_step_max,'// &
                'dvelgradavg_mag,cond_c066mod_max_sol,cond_c066mod_max_gas,cond_cgas_max,'// &
                'imin_sol_i,imin_sol_j,imin_sol_k,imin_gas_i,imin_gas_j,imin_gas_k'
            close (u)
        end if
        open (newunit=u, file=trim(path), status='old', position='append', action='write')
        write (u, '(*(g0,:,","))') iter, imacro, ipr, sim_time, wall_time, tdot, vals, imin
        flush (u)
        close (u)
    end subroutine write_diag

    function csv_path(dir, name) result(p)
        character(len=*), intent(in) :: dir, name
        character(len=512) :: p
        integer :: n
        n = len_trim(dir)
        if (n == 0) call fatal('output_dir is empty')
        if (dir(n:n) == '/') then
            p = trim(dir)//name
        else
            p = trim(dir)//'/'//name
        end if
    end function csv_path

    ! 3x3 row-major -> (11, 22, 33, 23, 13, 12)
    function v6(m) result(o)
        real(c_double), intent(in) :: m(9)
        real(8) :: o(6)
        o = [m(1), m(5), m(9), m(6), m(3), m(2)]
    end function v6

end module host_io
