! host_vtk.f90 -- legacy ASCII VTK field output (write_vtk_cell,
! lsevpfft/src/IO_functions.f90:945-1235): UNSTRUCTURED_GRID of hexahedra
! (cell type 12) on the current node coordinates, with the eleven CELL_DATA
! arrays gamdot, gamdot_acum, def_grad_p, stress, svm, def_grad, det_def_grad,
! gr_id, ph_id, epvm, velgrad; values below the smallest normal double are
! written as 0.
module host_vtk
    use iso_c_binding
    use creep_hip_iface
    use host_io
    implicit none
contains

    elemental real(8) function clampsub(a)
        real(8), intent(in) :: a
        clampsub = a
        if (abs(a) < tiny(1.0d0)) clampsub = 0.0d0
    end function clampsub

    subroutine fetch(ctx, name, buf, nc, nv)
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: name
        real(c_double), allocatable, intent(out) :: buf(:)
        integer, intent(out) :: nc
        integer(8), intent(in) :: nv
        integer(c_int) :: rc
        nc = cg_field_ncomp(ctx, trim(name)//c_null_char)
        if (nc < 1) call fatal('unknown field '//trim(name))
        if (trim(name) == 'xnode') then
            allocate (buf(nv))
        else
            allocate (buf(int(nc, 8) * nv))
        end if
        rc = cg_get_field(ctx, trim(name)//c_null_char, buf, int(size(buf), c_size_t))
        if (rc /= 0) call fatal('cg_get_field '//trim(name)//': '//trim(last_error(ctx)))
    end subroutine fetch

    subroutine write_vtk_cell(ctx, inp, istp, output_dir)
        type(c_ptr), intent(in) :: ctx
        type(input_t), intent(in) :: inp
        integer, intent(in) :: istp
        character(len=*), intent(in) :: output_dir
        integer :: n1, n2, n3, kx, ky, kz, u, nc, ncell, iph, nsmx, i
        integer(8) :: nv, np, v, ic
        character(len=5) :: str
        real(c_double), allocatable :: xn(:), f(:)
        integer, parameter :: colmaj(9) = [1, 4, 7, 2, 5, 8, 3, 6, 9]
        real(8) :: s(9), h, vm

        n1 = inp%npts1; n2 = inp%npts2; n3 = inp%npts3
        nv = int(n1, 8) * n2 * n3
        write (str, '(I5.5)') istp
        open (newunit=u, file=trim(output_dir)//'/microstr_cell_fields_stp_'//str//'.vtk', status='replace', &
              action='write')
        write (u, '(A)') '# vtk DataFile Version 3.0'
        write (u, '(A)') 'microstructure data'
        write (u, '(A)') 'ASCII'
        write (u, '(A)') 'DATASET UNSTRUCTURED_GRID'
        if (n3 == 1) then
            np = int(n1 + 1, 8) * (n2 + 1)
        else
            np = int(n1 + 1, 8) * (n2 + 1) * (n3 + 1)
        end if
        call fetch(ctx, 'xnode', xn, nc, 3 * np)
        write (u, '(A,I10,A)') 'POINTS', np, ' double'
        do ic = 0, np - 1
            write (u, '(3E18.8E3)') xn(3 * ic + 1), xn(3 * ic + 2), xn(3 * ic + 3)
        end do
        ncell = n1 * n2 * n3
        if (n3 == 1) then
            write (u, '(A,I10,I10)') 'CELLS', ncell, ncell * 5
            do ky = 1, n2
                do kx = 1, n1
                    write (u, '(5I10)') 4, nid(kx, ky, 1), nid(kx + 1, ky, 1), nid(kx, ky + 1, 1), nid(kx + 1, ky + 1, 1)
                end do
            end do
            write (u, '(A,I10)') 'CELL_TYPES', ncell
            do i = 1, ncell
                write (u, '(I5)') 8
            end do
        else
            write (u, '(A,I10,I10)') 'CELLS', ncell, ncell * 9
            do kz = 1, n3
                do ky = 1, n2
                    do kx = 1, n1
                        write (u, '(9I10)') 8, nid(kx, ky, kz), nid(kx + 1, ky, kz), nid(kx + 1, ky + 1, kz), &
                            nid(kx, ky + 1, kz), nid(kx, ky, kz + 1), nid(kx + 1, ky, kz + 1), &
                            nid(kx + 1, ky + 1, kz + 1), nid(kx, ky + 1, kz + 1)
                    end do
                end do
            end do
            write (u, '(A,I10)') 'CELL_TYPES', ncell
            do i = 1, ncell
                write (u, '(I5)') 12
            end do
        end if
        write (u, '(A,I10)') 'CELL_DATA', ncell
        write (u, '(A,I10)') 'FIELD FieldData', 11
        nsmx = 0
        do iph = 1, inp%nph
            nsmx = max(nsmx, inp%phase(iph)%nsyst)
        end do
        call fetch(ctx, 'gamdot', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'gamdot', nsmx, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(100E18.8E3)') (clampsub(f(nc * v + i)), i=1, nsmx)
        end do
        call fetch(ctx, 'gacumgr', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'gamdot_acum', 1, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(100E18.8E3)') clampsub(f(v + 1))
        end do
        call fetch(ctx, 'defgradp', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'def_grad_p', 9, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(9E18.8E3)') (clampsub(f(9 * v + colmaj(i))), i=1, 9)
        end do
        call fetch(ctx, 'sg', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'stress', 6, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(6E18.8E3)') clampsub(f(9 * v + 1)), clampsub(f(9 * v + 5)), clampsub(f(9 * v + 9)), &
                clampsub(f(9 * v + 2)), clampsub(f(9 * v + 6)), clampsub(f(9 * v + 3))
        end do
        write (u, '(A,I3,I10,A)') 'svm', 1, ncell, ' double'
        do v = 0, nv - 1
            s = f(9 * v + 1:9 * v + 9)
            h = (s(1) + s(5) + s(9)) / 3.d0
            vm = sqrt(1.5d0 * ((s(1) - h)**2 + (s(5) - h)**2 + (s(9) - h)**2 + s(2)**2 + s(3)**2 + s(6)**2 + &
                               s(4)**2 + s(7)**2 + s(8)**2))
            write (u, '(E18.8E3)') clampsub(vm)
        end do
        call fetch(ctx, 'defgrad', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'def_grad', 9, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(9E18.8E3)') (clampsub(f(9 * v + colmaj(i))), i=1, 9)
        end do
        call fetch(ctx, 'detF', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'det_def_grad', 1, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(E18.8E3)') clampsub(f(v + 1))
        end do
        write (u, '(A,I3,I10,A)') 'gr_id', 1, ncell, ' int'
        do v = 1, nv
            write (u, '(I10)') inp%grain(v)
        end do
        write (u, '(A,I3,I10,A)') 'ph_id', 1, ncell, ' int'
        do v = 1, nv
            write (u, '(I10)') inp%ph(v)
        end do
        call fetch(ctx, 'epvm', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'epvm', 1, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(E18.8E3)') clampsub(f(v + 1))
        end do
        call fetch(ctx, 'velgrad', f, nc, nv)
        write (u, '(A,I3,I10,A)') 'velgrad', 9, ncell, ' double'
        do v = 0, nv - 1
            write (u, '(9E18.8E3)') (clampsub(f(9 * v + colmaj(i))), i=1, 9)
        end do
        close (u)

    contains
        integer(8) function nid(i, j, k)
            integer, intent(in) :: i, j, k
            nid = (i - 1) + int(n1 + 1, 8) * ((j - 1) + int(n2 + 1, 8) * (k - 1))
        end function nid
    end subroutine write_vtk_cell

end module host_vtk
