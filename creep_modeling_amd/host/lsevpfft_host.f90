! lsevpfft_host.f90 -- drop-in for lsevpfft/LS-EVPFFT: the LS-EVPFFT driver
! (lsevpfft/src/LS-EVPFFT.f90:1-597) with every field operation on the GPU
! through libcreep_hip.so.
!
! Same command line (--in --out --nthreads --fftw_save --modal --mix, -h),
! same input files, same results.csv / diagnostics.csv rows and the same
! process / increment / fixed-point control flow: convergence test, Delta-t
! control with kinematic sub-stepping, rollback and ramp.  New flags:
! --precision fp64|mixed (spectral stage in FP32 for "mixed").
program lsevpfft_host
    use iso_c_binding
    use creep_hip_iface
    use host_io
    use host_vtk
    implicit none

    type(input_t) :: inp
    type(c_ptr) :: ctx
    type(cg_iter_state) :: st
    type(cg_iter_out) :: io
    type(cg_incr_out) :: ic
    integer(c_int) :: rc, nsub_c
    integer :: i, iph, ipr, imicro, imacroloop, nconverged, total_macrosteps, nsteps_left, nsteps_total_new
    integer :: nthreads, tmp_int, precision, nsub, last_iter, npass
    integer(8) :: nv, start, cnt_rate, now
    logical :: skip_next, modal, do_mix, did_converge, accept_step
    character(len=256) :: arg, fft_input_file, output_dir, fftw_save_path
    real(8) :: tdot, tdot_plan, proc_time_target, proc_time_done, remaining_time, sim_time, wall_time, tdot_used
    real(8) :: kin(4), vs(3, 3), h, evm, diagv(32), prev(6), cur(6)
    integer :: imin(6)
    real(8), allocatable :: cflat(:), dn(:), db(:), tau(:), th(:), hard(:)

    fft_input_file = 'fft.in'
    output_dir = 'evpfft_outputs/'
    fftw_save_path = ''
    nthreads = 0
    modal = .true.
    do_mix = .true.
    precision = 0
    skip_next = .false.
    write (*, *)
    do i = 1, command_argument_count()
        if (skip_next) then
            skip_next = .false.
            cycle
        end if
        call get_command_argument(i, arg)
        select case (arg)
        case ('--in')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            fft_input_file = trim(arg)
        case ('--out')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            output_dir = trim(arg)
        case ('--nthreads')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            read (arg, *) nthreads
        case ('--fftw_save')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            fftw_save_path = trim(arg)
        case ('--modal')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            read (arg, *) tmp_int
            modal = (tmp_int /= 0)
        case ('--mix')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            read (arg, *) tmp_int
            do_mix = (tmp_int /= 0)
        case ('--precision')
            skip_next = .true.
            call get_command_argument(i + 1, arg)
            if (trim(arg) == 'mixed') then
                precision = 1
            else if (trim(arg) == 'fp64') then
                precision = 0
            else
                write (*, '(3A)') 'Unknown precision "', trim(arg), '" (fp64|mixed)'
                stop
            end if
        case ('-h', '--help')
            write (*, '(A)') 'Optional inputs:'
            write (*, '(A)') '--in FILE | name of input file (fft.in by default)'
            write (*, '(A)') '--out DIR | output directory (evpfft_outputs by default)'
            write (*, '(A)') '--nthreads n | host threads (the field work runs on the GPU)'
            write (*, '(A)') '--fftw_save FILE | FFT shape tuning cache (read if it holds this grid, else written)'
            write (*, "(A)") '--modal 0-1 | toggle modified AL (default 1)'
            write (*, "(A)") '--mix 0-1 | velocity-gradient mixing (default 1)'
            write (*, "(A)") '--precision fp64|mixed | arithmetic of the FFT/Green stage (default fp64)'
            stop
        case default
            write (*, '(3A)') 'Unknown command-line option: "', trim(arg), '" use -h or --help for argument list'
            stop
        end select
    end do

    call execute_command_line('mkdir -p "'//trim(output_dir)//'"')
    call execute_command_line('find "'//trim(output_dir)//'" -maxdepth 1 -type f -name "*.vtk" -delete')
    call execute_command_line('find "'//trim(output_dir)//'" -maxdepth 1 -type f -name "*.out" -delete')
    if (len_trim(fftw_save_path) > 0) call execute_command_line('mkdir -p "$(dirname "'//trim(fftw_save_path)//'")"')

    call load_input(fft_input_file, inp)
    nv = int(inp%npts1, 8) * inp%npts2 * inp%npts3
    write (*, '(A,I0,A,I0,A,I0,A)') 'Grid (x,y,z)=(', inp%npts1, ',', inp%npts2, ',', inp%npts3, ') on the GPU'

    rc = cg_create(inp%npts1, inp%npts2, inp%npts3, inp%nph, inp%complgas, merge(1, 0, modal), merge(1, 0, do_mix), &
                   precision, ctx)
    if (rc /= 0) call fatal('cg_create failed: no usable GPU')
    do iph = 1, inp%nph
        if (inp%phase(iph)%igas) then
            allocate (cflat(81))
            cflat = 0.d0
            rc = cg_set_phase(ctx, iph, 1, 0, 0, [0], cflat, cflat, cflat, cflat, cflat, cflat, cflat, cflat, cflat)
            deallocate (cflat)
        else
            call set_phase(iph, inp%phase(iph))
        end if
        if (rc /= 0) call fatal('cg_set_phase: '//trim(last_error(ctx)))
    end do
    rc = cg_set_voxels(ctx, inp%euler, inp%grain, inp%ph)
    if (rc /= 0) call fatal('cg_set_voxels: '//trim(last_error(ctx)))
    if (inp%idispini > 0) then  ! init_defgrad (IO_functions.f90:152-154)
        rc = cg_set_displacement(ctx, inp%idispini, inp%dispini)
        if (rc /= 0) call fatal('cg_set_displacement: '//trim(last_error(ctx)))
    end if
    rc = cg_init(ctx, inp%proc(1))
    if (rc /= 0) call fatal_rc('cg_init: '//trim(last_error(ctx)), rc)
    ! the FFT shape tuning of the first pass is cached where the reference keeps its FFTW wisdom
    if (len_trim(fftw_save_path) > 0) then
        rc = cg_set_tuning_file(ctx, trim(fftw_save_path)//c_null_char)
        if (rc /= 0) call fatal('cg_set_tuning_file: '//trim(last_error(ctx)))
    end if
    if (have_thermo) call set_thermo()
    call write_vtk_cell(ctx, inp, 0, output_dir)

    call system_clock(start, cnt_rate)
    ! initial rows (LS-EVPFFT.f90:176-178): averages are zero except dvm = vm(dsim of process 1)
    ic = zero_incr()
    vs = reshape(inp%proc(1)%dsim, [3, 3])
    h = (vs(1, 1) + vs(2, 2) + vs(3, 3)) / 3.d0
    ic%dvmavg = sqrt(2.d0 / 3.d0 * ((vs(1, 1) - h)**2 + (vs(2, 2) - h)**2 + (vs(3, 3) - h)**2 + &
                                   vs(1, 2)**2 + vs(1, 3)**2 + vs(2, 3)**2 + vs(2, 1)**2 + vs(3, 1)**2 + vs(3, 2)**2))
    diagv = 0.d0
    imin = 0
    call write_diag(output_dir, .true., 0, 0, 0, 0.d0, 0.d0, 0.d0, diagv, imin)
    call write_results(output_dir, .true., 0, 0, 0, 0.d0, 0.d0, 0.d0, 0.d0, ic)

    total_macrosteps = sum(inp%proc(:)%nsteps)
    imacroloop = 1
    sim_time = 0.d0
    wall_time = 0.d0
    prev = 0.d0
    write (*, *)
    do ipr = 1, inp%nproc
        nconverged = 0
        tdot_plan = inp%proc(ipr)%tdot
        proc_time_target = dble(inp%proc(ipr)%nsteps) * tdot_plan
        proc_time_done = 0.d0
        imicro = 0
        do while (imicro < inp%proc(ipr)%nsteps)
            imicro = imicro + 1
            associate (p => inp%proc(ipr))
                if (p%ictrl == -1) then
                    remaining_time = max(0.d0, proc_time_target - proc_time_done)
                    if (remaining_time > 0.d0) then
                        p%tdot = min(tdot_plan, remaining_time)
                    else
                        p%tdot = tdot_plan
                    end if
                end if
                rc = cg_begin_increment(ctx, p, merge(1, 0, (imicro == 1 .and. ipr == 1) .or. p%iupdate == 1))
                if (rc /= 0) call fatal_rc('cg_begin_increment: '//trim(last_error(ctx)), rc)
                did_converge = .false.
                accept_step = .false.
                st%iter = 0
                st%mix = 1.d0
                st%imicro = imicro
                npass = 0
                do while (.not. (did_converge .or. accept_step))
                    call write_status(.false.)
                    rc = cg_iterate(ctx, p, st, io)
                    if (rc /= 0) call fatal_rc('cg_iterate: '//trim(last_error(ctx)), rc)
                    ! The reference's loop has no exit but convergence when ictrl /= -1 or
                    ! tdot > tdotmin (LS-EVPFFT.f90:361-429): an increment that never converges
                    ! would run forever; it is ended after 100 itmax + 1000 passes.
                    npass = npass + 1
                    if (npass > 100 * max(p%itmax, 1) + 1000) &
                        call fatal('fixed-point loop did not terminate (no convergence in 100 itmax + 1000 passes)')
                    if (p%ictrl == 0) then
                        vs = 0.5d0 * (reshape(io%velgradavg, [3, 3]) + transpose(reshape(io%velgradavg, [3, 3])))
                        h = (vs(1, 1) + vs(2, 2) + vs(3, 3)) / 3.d0
                        evm = sqrt(2.d0 / 3.d0 * ((vs(1, 1) - h)**2 + (vs(2, 2) - h)**2 + (vs(3, 3) - h)**2 + &
                                                 2.d0 * (vs(1, 2)**2 + vs(1, 3)**2 + vs(2, 3)**2)))
                        p%tdot = p%devmapp / evm
                        p%tdotref = p%tdot
                    end if
                    if (io%errs <= p%error .and. io%erre <= p%error .and. io%errsbc <= p%error) did_converge = .true.
                    if (p%ictrl == -1) then
                        if (p%tdot > p%tdotmin) then
                            if (.not. did_converge .and. st%iter == p%itmax) then
                                tdot = p%tdot
                                rc = cg_kinematic_substep(ctx, tdot, kin, nsub_c)
                                nsub = nsub_c
                                tdot = tdot / nsub
                                tdot = max(p%tdotmin, min(tdot, 0.5d0 * p%tdot))
                                nsub = ceiling(p%tdot / tdot)
                                if (nsub > 1) then
                                    call replan(p, tdot)
                                    nconverged = 0
                                    st%iter = 0
                                    rc = cg_rollback(ctx, p)
                                end if
                            else if (did_converge) then
                                tdot = p%tdot
                                rc = cg_kinematic_substep(ctx, tdot, kin, nsub_c)
                                nsub = nsub_c
                                tdot = tdot / nsub
                                if (kin(4) <= 1.0d-3) then
                                    tdot = max(p%tdotmin, min(tdot, 0.5d0 * p%tdot))
                                else
                                    tdot = max(p%tdotmin, tdot)
                                end if
                                nsub = ceiling(p%tdot / tdot)
                                if (nsub > 1) then
                                    call replan(p, tdot)
                                    did_converge = .false.
                                    nconverged = 0
                                    st%iter = 0
                                    st%mix = 1.d0
                                    rc = cg_rollback(ctx, p)
                                end if
                            end if
                        else if (st%iter == p%itmax) then
                            accept_step = .true.
                        end if
                    end if
                end do
                last_iter = st%iter
                rc = cg_end_increment(ctx, p, 1, ic)
                if (rc /= 0) call fatal_rc('cg_end_increment: '//trim(last_error(ctx)), rc)
                tdot_used = p%tdot
                if (p%ictrl == -1) then
                    proc_time_done = proc_time_done + tdot_used
                    p%tdot = tdot_plan
                end if
                ! field output cadence (LS-EVPFFT.f90:526-535)
                if (p%iwfields == 1 .and. p%iwstep /= 0) then
                    if (mod(imicro, abs(p%iwstep)) == 0 .or. (p%iwstep < 0 .and. imicro == p%nsteps)) &
                        call write_vtk_cell(ctx, inp, imacroloop, output_dir)
                end if
                sim_time = sim_time + tdot_used
                call system_clock(now)
                wall_time = dble((now - start) / cnt_rate)
                cur = [ic%detf_min_sol, ic%detf_min_gas, ic%volinc_max_sol, ic%volinc_max_gas, &
                       ic%stretchinc_max_sol, ic%stretchinc_max_gas]
                diagv = [ic%edotvmmx * tdot_used, p%tdotref, io%erre, io%errs, io%errsbc, p%scauchy(9), ic%sgavg(9), &
                         ic%sgsolavg(9), ic%detf_min_sol, ic%detf_min_gas, safe_rel(cur(1), prev(1)), &
                         safe_rel(cur(2), prev(2)), ic%volinc_max_sol, ic%volinc_max_gas, safe_rel(cur(3), prev(3)), &
                         safe_rel(cur(4), prev(4)), ic%stretchinc_max_sol, ic%stretchinc_max_gas, &
                         safe_rel(cur(5), prev(5)), safe_rel(cur(6), prev(6)), ic%rotang_max_sol, ic%rotang_max_gas, &
                         ic%dl_step_rms_sol, ic%dl_step_rms_gas, ic%dl_step_rms, ic%dl_step_max_sol, &
                         ic%dl_step_max_gas, ic%dl_step_max, ic%dvelgradavg_mag, ic%cond_c066mod_max_sol, &
                         ic%cond_c066mod_max_gas, ic%cond_cgas_max]
                prev = cur
                imin = [ic%imin_sol, ic%imin_gas]
                call write_diag(output_dir, .false., last_iter, imacroloop, ipr, sim_time, wall_time, tdot_used, diagv, imin)
                call write_results(output_dir, .false., last_iter, imacroloop, ipr, sim_time, wall_time, tdot_used, &
                                   p%tdotref, ic)
                ! ramp tdot back up (LS-EVPFFT.f90:557-581)
                if (p%ictrl == -1) then
                    if (tdot_plan < p%tdotref .and. nconverged >= 5) then
                        nconverged = 0
                        tdot_plan = min(2.0d0 * tdot_plan, p%tdotref)
                        remaining_time = max(0.0d0, proc_time_target - proc_time_done)
                        nsteps_left = ceiling(remaining_time / tdot_plan)
                        nsteps_total_new = imicro + nsteps_left
                        total_macrosteps = total_macrosteps + nsteps_total_new - p%nsteps
                        p%nsteps = nsteps_total_new
                        if (remaining_time > 0.0d0) then
                            p%tdot = min(tdot_plan, remaining_time)
                        else
                            p%tdot = tdot_plan
                        end if
                    end if
                else if (p%tdot < p%tdotref .and. nconverged >= 5) then
                    nconverged = 0
                    tdot = p%tdot
                    p%tdot = min(2.0 * p%tdot, p%tdotref)
                    nsteps_left = p%nsteps - imicro
                    nsteps_left = ceiling(dble(nsteps_left) * tdot / p%tdot)
                    total_macrosteps = total_macrosteps + nsteps_left - (p%nsteps - imicro)
                    p%nsteps = imicro + nsteps_left
                end if
                if (did_converge) nconverged = nconverged + 1
                rc = cg_calc_c0(ctx, p)
                if (rc /= 0) call fatal_rc('cg_calc_c0: '//trim(last_error(ctx)), rc)
            end associate
            imacroloop = imacroloop + 1
        end do
    end do
    call write_status(.true.)
    write (*, *)
    rc = cg_destroy(ctx)

contains

    subroutine set_phase(iph, p)
        integer, intent(in) :: iph
        type(phase_t), intent(in) :: p
        integer :: ns, i, j, k, l, is
        ns = p%nsyst
        allocate (cflat(81), dn(3 * ns), db(3 * ns), tau(3 * ns), th(2 * ns), hard(ns * ns))
        do i = 1, 3
            do j = 1, 3
                do k = 1, 3
                    do l = 1, 3
                        cflat(27 * (i - 1) + 9 * (j - 1) + 3 * (k - 1) + l) = p%c(i, j, k, l)
                    end do
                end do
            end do
        end do
        do is = 1, ns
            dn(3 * (is - 1) + 1:3 * is) = p%dnca(:, is)
            db(3 * (is - 1) + 1:3 * is) = p%dbca(:, is)
            tau(3 * (is - 1) + 1:3 * is) = p%tau(is, :)
            th(2 * (is - 1) + 1) = p%thet(is, 0)
            th(2 * (is - 1) + 2) = p%thet(is, 1)
            do j = 1, ns
                hard(ns * (is - 1) + j) = p%hard(is, j)
            end do
        end do
        rc = cg_set_phase(ctx, iph, 0, ns, p%nmodes, p%nsm, dn, db, tau, th, p%nrs, p%gamd0, p%hpfac, hard, cflat)
        deallocate (cflat, dn, db, tau, th, hard)
    end subroutine set_phase

    ! per-voxel eigenstrain rates from the grain table (micro%voxel%eth = eth(:,:,grain))
    subroutine set_thermo()
        real(c_double), allocatable :: e(:)
        integer(8) :: v
        integer :: g, r, c
        allocate (e(9 * nv))
        e = 0.d0
        do v = 1, nv
            g = inp%grain(v)
            if (g < 1 .or. g > size(eth_grain, 3)) call fatal('grain id without an eigenstrain rate in the ithermo file')
            do r = 1, 3
                do c = 1, 3
                    e(9 * (v - 1) + 3 * (r - 1) + c) = eth_grain(r, c, g)
                end do
            end do
        end do
        rc = cg_set_thermo(ctx, e)
        if (rc /= 0) call fatal('cg_set_thermo: '//trim(last_error(ctx)))
        deallocate (e)
    end subroutine set_thermo

    subroutine replan(p, tdot_new)
        type(cg_process), intent(inout) :: p
        real(8), intent(in) :: tdot_new
        remaining_time = max(0.0d0, proc_time_target - proc_time_done)
        tdot_plan = tdot_new
        nsteps_total_new = (imicro - 1) + ceiling(remaining_time / tdot_plan)
        total_macrosteps = total_macrosteps + nsteps_total_new - p%nsteps
        p%nsteps = nsteps_total_new
        p%tdot = min(tdot_plan, remaining_time)
    end subroutine replan

    pure real(8) function safe_rel(x, xp)
        real(8), intent(in) :: x, xp
        if (xp <= tiny(1.0d0) .or. xp >= huge(1.0d0) * 0.5d0) then
            safe_rel = 0.0d0
        else
            safe_rel = abs(x - xp) / xp
        end if
    end function safe_rel

    function zero_incr() result(z)
        type(cg_incr_out) :: z
        z%eavg = 0; z%sgsolavg = 0; z%epavg = 0; z%eelavg = 0; z%edotpavg = 0; z%velgradavg = 0; z%sgavg = 0
        z%evmavg = 0; z%evmpavg = 0; z%dvmavg = 0; z%dvmpavg = 0; z%svmavg = 0; z%lvmavg = 0; z%edotvmmx = 0
        z%gammaavg = 0; z%gavg = 0; z%geffavg = 0
    end function zero_incr

    ! progress line (various_functions.f90:2693-2743)
    subroutine write_status(final)
        logical, intent(in) :: final
        character(len=256) :: status
        integer :: pct, im
        im = imacroloop
        pct = (100 * imacroloop) / max(1, total_macrosteps)
        if (final) then
            im = imacroloop - 1
            pct = 100
        end if
        write (status, '(A,I0,A,I0,A,I0,A,I0,A,I0,A,I0,A,I0,A,I0,A)') 'MACROSTEP=', im, ' SimTime=', nint(sim_time), &
            's Wall=', nint(wall_time), 's | PROC=', ipr, ' STEP=', imicro, ' TDOT=', &
            nint(inp%proc(min(max(ipr, 1), inp%nproc))%tdot), 's ITER=', st%iter + 1, ' | Completion: ', pct, '%'
        write (*, '(A)', advance='no') achar(13)//achar(27)//'[2K'//achar(13)//trim(status)
        flush (6)
    end subroutine write_status

end program lsevpfft_host
