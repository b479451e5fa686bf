! creep_hip_iface.f90 -- ISO_C_BINDING interface to libcreep_hip.so
! (declarations mirror include/creep_hip.h one for one).
module creep_hip_iface
    use iso_c_binding
    implicit none

    ! process block of fft.in (types.f90:120-152 fields used by the loop);
    ! 3x3 tensors row-major: a(3*(i-1)+j) = A(i,j)
    type, bind(C) :: cg_process
        integer(c_int) :: iudot(9)
        real(c_double) :: udot(9)
        integer(c_int) :: idsim(6), iscau(6)
        real(c_double) :: scauchy(9), dsim(9)
        real(c_double) :: tdot, tdotref, tdotmin, devmmx, devmapp
        integer(c_int) :: ictrl, nsteps
        real(c_double) :: error, erroral
        integer(c_int) :: itmax, iupdate, iuphard, iwfields, iwstep
        real(c_double) :: xc0
        integer(c_int) :: igamma, ithermo
    end type cg_process

    type, bind(C) :: cg_iter_state
        integer(c_int) :: iter
        real(c_double) :: mix
        integer(c_int) :: imicro
    end type cg_iter_state

    type, bind(C) :: cg_iter_out
        real(c_double) :: errs, erre, errsbc
        real(c_double) :: velgradavg(9), sgavg(9), sgavgsol(9), dvelgradavg(9)
        real(c_double) :: svmavg, dl_iter
        integer(c_int64_t) :: nr_steps, evals, outer
    end type cg_iter_out

    type, bind(C) :: cg_incr_out
        real(c_double) :: eavg(9), sgsolavg(9), epavg(9), eelavg(9), edotpavg(9), velgradavg(9), sgavg(9)
        real(c_double) :: evmavg, evmpavg, dvmavg, dvmpavg, svmavg, lvmavg, edotvmmx
        real(c_double) :: gammaavg, gavg, geffavg
        real(c_double) :: detf_min, detf_min_sol, detf_min_gas
        real(c_double) :: volinc_max_sol, volinc_max_gas, stretchinc_max_sol, stretchinc_max_gas
        real(c_double) :: rotang_max_sol, rotang_max_gas
        real(c_double) :: dl_step_rms_sol, dl_step_rms_gas, dl_step_rms, dl_step_max_sol, dl_step_max_gas, dl_step_max
        real(c_double) :: dvelgradavg_mag, cond_c066mod_max_sol, cond_c066mod_max_gas, cond_cgas_max
        integer(c_int) :: imin_sol(3), imin_gas(3)
    end type cg_incr_out

    interface
        integer(c_int) function cg_create(n1, n2, n3, nph, complgas, modal, do_mix, precision, ctx) bind(C, name='cg_create')
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: n1, n2, n3, nph, modal, do_mix, precision
            real(c_double), value :: complgas
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function cg_destroy(ctx) bind(C, name='cg_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function cg_last_error(ctx) bind(C, name='cg_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function cg_set_phase(ctx, iph, igas, nsyst, nmodes, nsm, dnca, dbca, tau, thet, nrs, gamd0, &
                                             hpfac, hard, c) bind(C, name='cg_set_phase')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: iph, igas, nsyst, nmodes
            integer(c_int) :: nsm(*)
            real(c_double) :: dnca(*), dbca(*), tau(*), thet(*), nrs(*), gamd0(*), hpfac(*), hard(*), c(*)
        end function
        integer(c_int) function cg_set_voxels(ctx, euler, grain, phase) bind(C, name='cg_set_voxels')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double) :: euler(*)
            integer(c_int) :: grain(*), phase(*)
        end function
        integer(c_int) function cg_set_displacement(ctx, idispini, disp) bind(C, name='cg_set_displacement')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: idispini
            real(c_double) :: disp(*)
        end function
        integer(c_int) function cg_init(ctx, p1) bind(C, name='cg_init')
            import :: c_int, c_ptr, cg_process
            type(c_ptr), value :: ctx
            type(cg_process) :: p1
        end function
        integer(c_int) function cg_begin_increment(ctx, p, do_schmid) bind(C, name='cg_begin_increment')
            import :: c_int, c_ptr, cg_process
            type(c_ptr), value :: ctx
            type(cg_process) :: p
            integer(c_int), value :: do_schmid
        end function
        integer(c_int) function cg_iterate(ctx, p, st, out) bind(C, name='cg_iterate')
            import :: c_int, c_ptr, cg_process, cg_iter_state, cg_iter_out
            type(c_ptr), value :: ctx
            type(cg_process) :: p
            type(cg_iter_state) :: st
            type(cg_iter_out) :: out
        end function
        integer(c_int) function cg_rollback(ctx, p) bind(C, name='cg_rollback')
            import :: c_int, c_ptr, cg_process
            type(c_ptr), value :: ctx
            type(cg_process) :: p
        end function
        integer(c_int) function cg_kinematic_substep(ctx, tdot, out4, nsub) bind(C, name='cg_kinematic_substep')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: tdot
            real(c_double) :: out4(4)
            integer(c_int) :: nsub
        end function
        integer(c_int) function cg_set_tuning_file(ctx, path) bind(C, name='cg_set_tuning_file')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char) :: path(*)
        end function
        integer(c_int) function cg_set_thermo(ctx, eth) bind(C, name='cg_set_thermo')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double) :: eth(*)
        end function
        ! device pointers in and out, nothing allocated or synchronised (callers that cannot allocate)
        integer(c_int) function cg_kinematic_substep_noalloc(nvox, velgrad, ld, detf, tdot, out5, stream) &
            bind(C, name='cg_kinematic_substep_noalloc')
            import :: c_int, c_long, c_double, c_ptr
            integer(c_long), value :: nvox, ld
            type(c_ptr), value :: velgrad, detf, out5, stream
            real(c_double), value :: tdot
        end function
        integer(c_int) function cg_end_increment(ctx, p, scan, out) bind(C, name='cg_end_increment')
            import :: c_int, c_ptr, cg_process, cg_incr_out
            type(c_ptr), value :: ctx
            type(cg_process) :: p
            integer(c_int), value :: scan
            type(cg_incr_out) :: out
        end function
        integer(c_int) function cg_calc_c0(ctx, p) bind(C, name='cg_calc_c0')
            import :: c_int, c_ptr, cg_process
            type(c_ptr), value :: ctx
            type(cg_process) :: p
        end function
        integer(c_int) function cg_field_ncomp(ctx, name) bind(C, name='cg_field_ncomp')
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char) :: name(*)
        end function
        integer(c_int) function cg_get_field(ctx, name, out, n) bind(C, name='cg_get_field')
            import :: c_int, c_ptr, c_char, c_double, c_size_t
            type(c_ptr), value :: ctx
            character(kind=c_char) :: name(*)
            real(c_double) :: out(*)
            integer(c_size_t), value :: n
        end function
    end interface

contains

    ! error text of the last failing call
    function last_error(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: i
        msg = ''
        p = cg_last_error(ctx)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [512])
        do i = 1, 512
            if (s(i) == c_null_char) exit
            msg(i:i) = s(i)
        end do
    end function last_error

end module creep_hip_iface
