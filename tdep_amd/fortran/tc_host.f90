! Fortran host of the three-phonon thermal-conductivity path on one MI355X.
!
! Takes the place of src/thermal_conductivity/main.f90 for the part of the program built
! here: reads infile.ucposcar, infile.forceconstant, infile.forceconstant_thirdorder and, with
! --readiso, infile.isotopes from the working directory (formats of
! type_crystalstructure_io.f90:9-350, :921-959, type_forceconstant_secondorder_io.f90:109-247,
! type_forceconstant_thirdorder.f90:243-311), hands everything numeric to libtc3 through the
! iso_c_binding interface (module tc3_api), and writes outfile.thermal_conductivity
! (main.f90:246-296), outfile.thermal_conductivity_grid.hdf5 (main.f90:323-328) and
! outfile.cumulative_thermal_conductivity.hdf5 (main.f90:298-340) for every requested
! temperature (files of temperatures after the first get a ".T<temperature>"
! suffix: all temperatures share one scattering pass).
!
! Options (options.f90:44-237, the subset of this path):
!   -qg n1 n2 n3 | --qpoint_grid n1 n2 n3     q-mesh (26 26 26)
!   -qg3ph n1 n2 n3 | --qpoint_grid3ph ...     Monte-Carlo q2 grid of the three-phonon term (mesh)
!   -qg4ph n1 n2 n3 | --qpoint_grid4ph ...     Monte-Carlo grids of the four-phonon term (mesh)
!   --seed s                                   seed of the Monte-Carlo draws (needed with -qg3ph, -qg4ph)
!   --temperature T [T ...]                    temperatures in K (300)
!   -it 1|2|6 | --integrationtype 1|2|6        fixed, adaptive or velocity-difference Gaussian (2)
!   --sigma s                                  fixed smearing in THz, integrationtype 1; also
!                                              scales the cumulative kappa's smoothing (1.0)
!   --max_mfp L                                boundary scattering, mean free path limit in m (off)
!   --mfppts n, --freqpts n                    points of the mean-free-path / frequency axes (1000)
!   --dosintegrationtype 1|2|3                 Gaussian, adaptive Gaussian, tetrahedron spectra (3)
!   --dossigma s                               smearing scale of the spectra (1.0)
!   --unit thz|mev|icm                         frequency unit of the spectra (thz)
!   --iterative_tolerance x                    BTE tolerance (1e-5)
!   --iterative_maxsteps n                     BTE iterations (200)
!   --classical, --noisotope, --notr
!   --readiso                                  masses from infile.isotopes (else natural isotopes)
!   --nothirdorder                             no three-phonon term (isotope, four-phonon, boundary only)
!   --nosym                                    mesh reduced by the identity and time reversal only
!   --precision fp64|fp32|bf16x3               precision of the |Psi|^2 contraction (fp64)
!   --no-offdiag                               skip the off-diagonal (coherence) term
!   --fourthorder                              add four-phonon scattering from
!                                              infile.forceconstant_fourthorder (full q-grid)
!   --device n                                 HIP device (0)
!   --inputs-only                              read the inputs, build the mesh on the host and
!                                              print sizes and input checksums; no device work
module tc_host_io
    use iso_c_binding
    use tc3_api, only: tc3_natural_isotopes
    implicit none
    private
    public :: fatal, read_ucposcar, read_isotopes, natural_masses, read_fc2, read_fc3, read_fc4, write_outfile, temperature_suffix

    ! konstanter.f90:36-91
    real(c_double), parameter :: bohr_to_a = 0.52917721067_c_double
    real(c_double), parameter :: a_to_bohr = 1.0_c_double/bohr_to_a
    real(c_double), parameter :: ev_to_hartree = 1.0_c_double/27.21138602_c_double
    real(c_double), parameter :: amu_to_emu = 1.660539040e-27_c_double/9.10938356e-31_c_double
    real(c_double), parameter :: sqtol = 1.0e-10_c_double
    ! gottochblandat.f90:1286-1374: magnitudes lo_chop snaps to
    real(c_double), parameter :: snapv(34) = [0.1_c_double, 0.1111111111111111_c_double, 0.125_c_double, &
        0.1428571428571428_c_double, 0.1666666666666667_c_double, 0.2_c_double, 0.2222222222222222_c_double, &
        0.25_c_double, 0.2857142857142857_c_double, 0.3_c_double, 0.3333333333333333_c_double, 0.375_c_double, &
        0.4_c_double, 0.4285714285714285_c_double, 0.4444444444444444_c_double, 0.5_c_double, &
        0.5555555555555556_c_double, 0.5714285714285714_c_double, 0.6_c_double, 0.625_c_double, &
        0.6666666666666667_c_double, 0.7_c_double, 0.7071067811865475_c_double, 0.7142857142857143_c_double, &
        0.75_c_double, 0.7777777777777778_c_double, 0.8_c_double, 0.8333333333333334_c_double, &
        0.8571428571428571_c_double, 0.8660254037844386_c_double, 0.875_c_double, 0.8888888888888888_c_double, &
        0.9_c_double, 1.0_c_double]
    character(len=2), parameter :: elements(103) = [character(len=2) :: &
        'H ', 'He', 'Li', 'Be', 'B ', 'C ', 'N ', 'O ', 'F ', 'Ne', 'Na', 'Mg', 'Al', 'Si', 'P ', 'S ', 'Cl', 'Ar', &
        'K ', 'Ca', 'Sc', 'Ti', 'V ', 'Cr', 'Mn', 'Fe', 'Co', 'Ni', 'Cu', 'Zn', 'Ga', 'Ge', 'As', 'Se', 'Br', 'Kr', &
        'Rb', 'Sr', 'Y ', 'Zr', 'Nb', 'Mo', 'Tc', 'Ru', 'Rh', 'Pd', 'Ag', 'Cd', 'In', 'Sn', 'Sb', 'Te', 'I ', 'Xe', &
        'Cs', 'Ba', 'La', 'Ce', 'Pr', 'Nd', 'Pm', 'Sm', 'Eu', 'Gd', 'Tb', 'Dy', 'Ho', 'Er', 'Tm', 'Yb', 'Lu', 'Hf', &
        'Ta', 'W ', 'Re', 'Os', 'Ir', 'Pt', 'Au', 'Hg', 'Tl', 'Pb', 'Bi', 'Po', 'At', 'Rn', 'Fr', 'Ra', 'Ac', 'Th', &
        'Pa', 'U ', 'Np', 'Pu', 'Am', 'Cm', 'Bk', 'Cf', 'Es', 'Fm', 'Md', 'No', 'Lr']

contains

    subroutine fatal(code, msg)
        integer, intent(in) :: code
        character(len=*), intent(in) :: msg
        write (0, '(A)') 'tc_host: '//trim(msg)
        select case (code)
        case (1); error stop 1
        case (2); error stop 2
        case (3); error stop 3
        case (4); error stop 4
        case (5); error stop 5
        case (6); error stop 6
        case default; error stop 7
        end select
    end subroutine

    ! lo_chop: snap x to 0 or a listed simple value when within tol.
    elemental function chop(x, tol) result(y)
        real(c_double), intent(in) :: x, tol
        real(c_double) :: y
        integer :: i
        y = x
        if (abs(x) < tol) then
            y = 0.0_c_double
            return
        end if
        do i = 1, size(snapv)
            if (abs(x - snapv(i)) < tol) then
                y = snapv(i)
                return
            end if
        end do
        do i = 1, size(snapv)
            if (abs(x + snapv(i)) < tol) then
                y = -snapv(i)
                return
            end if
        end do
    end function

    ! fractional coordinate in [0, 1)
    elemental function clean_fractional(x0) result(x)
        real(c_double), intent(in) :: x0
        real(c_double) :: x
        x = x0
        do
            if (abs(x) < sqtol) x = 0.0_c_double
            if (abs(x - 1.0_c_double) < sqtol) x = 0.0_c_double
            if (x > 1.0_c_double) x = x - 1.0_c_double
            if (x < 0.0_c_double) x = x + 1.0_c_double
            if (x > -sqtol .and. x < 1.0_c_double) exit
        end do
    end function

    function inverse3(m) result(r)
        real(c_double), intent(in) :: m(3, 3)
        real(c_double) :: r(3, 3), det
        det = m(1, 1)*(m(2, 2)*m(3, 3) - m(2, 3)*m(3, 2)) - m(1, 2)*(m(2, 1)*m(3, 3) - m(2, 3)*m(3, 1)) + &
              m(1, 3)*(m(2, 1)*m(3, 2) - m(2, 2)*m(3, 1))
        r(1, 1) = m(2, 2)*m(3, 3) - m(2, 3)*m(3, 2)
        r(1, 2) = m(1, 3)*m(3, 2) - m(1, 2)*m(3, 3)
        r(1, 3) = m(1, 2)*m(2, 3) - m(1, 3)*m(2, 2)
        r(2, 1) = m(2, 3)*m(3, 1) - m(2, 1)*m(3, 3)
        r(2, 2) = m(1, 1)*m(3, 3) - m(1, 3)*m(3, 1)
        r(2, 3) = m(1, 3)*m(2, 1) - m(1, 1)*m(2, 3)
        r(3, 1) = m(2, 1)*m(3, 2) - m(2, 2)*m(3, 1)
        r(3, 2) = m(1, 2)*m(3, 1) - m(1, 1)*m(3, 2)
        r(3, 3) = m(1, 1)*m(2, 2) - m(1, 2)*m(2, 1)
        r = r/det
    end function

    integer function count_words(line)
        character(len=*), intent(in) :: line
        integer :: i
        logical :: inword
        count_words = 0
        inword = .false.
        do i = 1, len_trim(line)
            if (line(i:i) /= ' ' .and. line(i:i) /= char(9)) then
                if (.not. inword) count_words = count_words + 1
                inword = .true.
            else
                inword = .false.
            end if
        end do
    end function

    subroutine open_input(u, path)
        integer, intent(out) :: u
        character(len=*), intent(in) :: path
        integer :: ios
        open (newunit=u, file=path, status='old', action='read', iostat=ios)
        if (ios /= 0) call fatal(6, 'could not open "'//path//'"')
    end subroutine

    ! POSCAR-style crystal structure.  latvec(:, i) is lattice vector i in Bohr, rfrac(:, a)
    ! fractional, species(a) 0-based id by first appearance of the element.
    subroutine read_ucposcar(path, na, latvec, rfrac, species, elist)
        character(len=*), intent(in) :: path
        integer, intent(out) :: na
        real(c_double), intent(out) :: latvec(3, 3)
        real(c_double), allocatable, intent(out) :: rfrac(:, :)
        integer(c_int), allocatable, intent(out) :: species(:)
        character(len=:), allocatable, intent(out) :: elist   ! elements in order, space separated
        character(len=1024) :: line
        character(len=8), allocatable :: sym(:), usym(:)
        integer, allocatable :: counts(:), z(:)
        real(c_double) :: scale, m(3, 3), im(3, 3), v(3), det
        integer :: u, nsp, i, j, a, nu
        character :: mode
        call open_input(u, path)
        read (u, '(A)') line
        read (u, *) scale
        do i = 1, 3
            read (u, *) m(:, i)
        end do
        read (u, '(A)') line
        nsp = count_words(line)
        allocate (sym(nsp), counts(nsp))
        read (line, *) sym
        read (u, *) counts
        read (u, '(A)') line
        line = adjustl(line)
        if (line(1:1) == 's' .or. line(1:1) == 'S') then
            read (u, '(A)') line
            line = adjustl(line)
        end if
        mode = line(1:1)
        if (index('CcDd', mode) == 0 .or. mode == ' ') &
            call fatal(3, path//': specify either Cartesian or direct coordinates')
        if (scale > 0) then
            m = m*scale
        else
            det = m(1, 1)*(m(2, 2)*m(3, 3) - m(2, 3)*m(3, 2)) - m(1, 2)*(m(2, 1)*m(3, 3) - m(2, 3)*m(3, 1)) + &
                  m(1, 3)*(m(2, 1)*m(3, 2) - m(2, 2)*m(3, 1))
            m = m*(abs(scale)/abs(det))**(1.0_c_double/3.0_c_double)
        end if
        na = sum(counts)
        allocate (rfrac(3, na), species(na), z(na), usym(nsp))
        a = 0
        do i = 1, nsp
            j = findloc(elements, trim(sym(i)), dim=1)
            if (j == 0) call fatal(3, path//': unknown element "'//trim(sym(i))//'"')
            z(a + 1:a + counts(i)) = j
            a = a + counts(i)
        end do
        im = inverse3(m)
        do a = 1, na
            read (u, *) v
            if (mode == 'C' .or. mode == 'c') v = matmul(im, v)
            rfrac(:, a) = chop(clean_fractional(clean_fractional(v)), sqtol)
        end do
        close (u)
        nu = 0
        do a = 1, na
            j = 0
            do i = 1, nu
                if (usym(i) == elements(z(a))) j = i
            end do
            if (j == 0) then
                nu = nu + 1
                usym(nu) = elements(z(a))
                j = nu
            end if
            species(a) = j - 1
        end do
        elist = trim(usym(1))
        do i = 2, nu
            elist = elist//' '//trim(usym(i))
        end do
        latvec = chop(m*a_to_bohr, sqtol)
    end subroutine

    ! Masses (emu) as the concentration-weighted mean and the mass-disorder parameter
    ! g = sum_i c_i (1 - m_i/m)^2 (type_crystalstructure.f90:1505-1519).
    subroutine read_isotopes(path, na, mass, disorder)
        character(len=*), intent(in) :: path
        integer, intent(in) :: na
        real(c_double), intent(out) :: mass(na), disorder(na)
        real(c_double), allocatable :: c(:), mm(:)
        integer :: u, a, n, j
        call open_input(u, path)
        do a = 1, na
            read (u, *) n
            allocate (c(n), mm(n))
            do j = 1, n
                read (u, *) c(j), mm(j)
            end do
            call distribution(c, mm, mass(a), disorder(a))
            deallocate (c, mm)
        end do
        close (u)
    end subroutine

    ! Masses and mass-disorder parameters from the natural isotope distribution of every atom's
    ! element (naturaldistribution, type_crystalstructure.f90:1521-2466, through
    ! tc3_natural_isotopes), normalised and averaged as read_isotopes does.  elist holds the
    ! species symbols separated by blanks, species(a) the 0-based species of atom a.
    subroutine natural_masses(na, species, elist, mass, disorder)
        integer, intent(in) :: na
        integer(c_int), intent(in) :: species(na)
        character(len=*), intent(in) :: elist
        real(c_double), intent(out) :: mass(na), disorder(na)
        character(len=8) :: sym(na)
        real(c_double) :: c(16), mm(16)
        integer(c_int) :: n
        integer :: a, ns, i0, i1
        ns = 0
        i0 = 1
        do while (i0 <= len_trim(elist))
            if (elist(i0:i0) == ' ') then
                i0 = i0 + 1
                cycle
            end if
            i1 = i0
            do while (i1 < len_trim(elist))
                if (elist(i1 + 1:i1 + 1) == ' ') exit
                i1 = i1 + 1
            end do
            ns = ns + 1
            sym(ns) = elist(i0:i1)
            i0 = i1 + 1
        end do
        do a = 1, na
            if (tc3_natural_isotopes(trim(sym(species(a) + 1))//c_null_char, 16_c_int, c, mm, n) /= 0) &
                call fatal(5, 'no isotope distributions available for '//trim(sym(species(a) + 1))// &
                           ', are you sure it is a stable element?')
            call distribution(c(1:n), mm(1:n), mass(a), disorder(a))
        end do
    end subroutine

    ! Mean mass (emu) and g = sum_i c_i (m_i - m)^2 / m^2 of abundances c and masses mm (amu),
    ! abundances normalised first (type_crystalstructure_io.f90:953-957).
    subroutine distribution(c0, mm0, mass, disorder)
        real(c_double), intent(in) :: c0(:), mm0(:)
        real(c_double), intent(out) :: mass, disorder
        real(c_double) :: c(size(c0)), mm(size(c0)), tot
        integer :: j
        mm = mm0*amu_to_emu
        tot = 0.0_c_double
        do j = 1, size(c0)
            tot = tot + c0(j)
        end do
        c = c0/tot
        mass = 0.0_c_double
        do j = 1, size(c0)
            mass = mass + c(j)*mm(j)
        end do
        disorder = 0.0_c_double
        do j = 1, size(c0)
            disorder = disorder + c(j)*(mm(j) - mass)**2
        end do
        disorder = disorder/mass**2
    end subroutine

    ! Second-order force constants: pair list with Cartesian lattice vectors (Bohr) and
    ! m(:, :, p) in Ha/Bohr^2; atoms 0-based.
    subroutine read_fc2(path, na, latvec, np, a1, a2, lv, m)
        character(len=*), intent(in) :: path
        integer, intent(in) :: na
        real(c_double), intent(in) :: latvec(3, 3)
        integer, intent(out) :: np
        integer(c_int), allocatable, intent(out) :: a1(:), a2(:)
        real(c_double), allocatable, intent(out) :: lv(:, :), m(:, :, :)
        integer :: u, n0, a, n, k, p, polar, j
        real(c_double) :: rc, v(3)
        integer, allocatable :: nper(:)
        call open_input(u, path)
        read (u, *) n0
        if (n0 /= na) call fatal(1, 'different number of atoms in "'//path//'" and the structure')
        read (u, *) rc
        ! first pass: counts
        allocate (nper(na))
        do a = 1, na
            read (u, *) nper(a)
            do k = 1, nper(a)
                do j = 1, 5
                    read (u, *)
                end do
            end do
        end do
        read (u, *) polar
        if (polar /= 0) call fatal(5, 'polar force constants (LO-TO) are not part of this path')
        rewind (u)
        read (u, *) n0
        read (u, *) rc
        np = sum(nper)
        allocate (a1(np), a2(np), lv(3, np), m(3, 3, np))
        p = 0
        do a = 1, na
            read (u, *) n
            do k = 1, n
                p = p + 1
                a1(p) = a - 1
                read (u, *) a2(p)
                a2(p) = a2(p) - 1
                read (u, *) v
                lv(:, p) = chop(matmul(latvec, v), 1.0e-13_c_double)
                do j = 1, 3
                    read (u, *) m(j, :, p)
                end do
            end do
        end do
        close (u)
        m = m*ev_to_hartree/a_to_bohr**2
    end subroutine

    ! Third-order force constants: triplet list, Cartesian lattice vectors of atoms 2 and 3
    ! (Bohr), m(:, :, :, t) in Ha/Bohr^3 unweighted; atoms 0-based.
    ! type_forceconstant_fourthorder.f90:259-299: per atom a count, then per quartet four atom
    ! indices, four fractional lattice vectors and 27 rows of the tensor (eV/A^4)
    subroutine read_fc4(path, na, latvec, nq, a1, a2, a3, a4, lv2, lv3, lv4, m)
        character(len=*), intent(in) :: path
        integer, intent(in) :: na
        real(c_double), intent(in) :: latvec(3, 3)
        integer, intent(out) :: nq
        integer(c_int), allocatable, intent(out) :: a1(:), a2(:), a3(:), a4(:)
        real(c_double), allocatable, intent(out) :: lv2(:, :), lv3(:, :), lv4(:, :), m(:, :, :, :, :)
        integer :: u, n0, a, k, t, i, j, l, idum, ncount
        real(c_double) :: rc, v(3)
        integer, allocatable :: nper(:)
        call open_input(u, path)
        read (u, *) n0
        if (n0 /= na) call fatal(1, 'different number of atoms in "'//path//'" and the structure')
        read (u, *) rc
        allocate (nper(na))
        do a = 1, na
            read (u, *) nper(a)
            do k = 1, nper(a)
                do j = 1, 35
                    read (u, *)
                end do
            end do
        end do
        rewind (u)
        read (u, *) n0
        read (u, *) rc
        nq = sum(nper)
        allocate (a1(nq), a2(nq), a3(nq), a4(nq), lv2(3, nq), lv3(3, nq), lv4(3, nq), m(3, 3, 3, 3, nq))
        t = 0
        do a = 1, na
            read (u, *) ncount
            do k = 1, ncount
                t = t + 1
                a1(t) = a - 1
                read (u, *) idum
                read (u, *) a2(t)
                read (u, *) a3(t)
                read (u, *) a4(t)
                a2(t) = a2(t) - 1
                a3(t) = a3(t) - 1
                a4(t) = a4(t) - 1
                read (u, *) v
                read (u, *) v
                lv2(:, t) = chop(matmul(latvec, v), sqtol)
                read (u, *) v
                lv3(:, t) = chop(matmul(latvec, v), sqtol)
                read (u, *) v
                lv4(:, t) = chop(matmul(latvec, v), sqtol)
                do i = 1, 3
                    do j = 1, 3
                        do l = 1, 3
                            read (u, *) m(i, j, l, :, t)
                        end do
                    end do
                end do
            end do
        end do
        close (u)
        m = m*ev_to_hartree/a_to_bohr**4
    end subroutine

    subroutine read_fc3(path, na, latvec, nt, a1, a2, a3, lv2, lv3, m)
        character(len=*), intent(in) :: path
        integer, intent(in) :: na
        real(c_double), intent(in) :: latvec(3, 3)
        integer, intent(out) :: nt
        integer(c_int), allocatable, intent(out) :: a1(:), a2(:), a3(:)
        real(c_double), allocatable, intent(out) :: lv2(:, :), lv3(:, :), m(:, :, :, :)
        integer :: u, n0, a, k, t, i, j, idum, ncount
        real(c_double) :: rc, v(3)
        integer, allocatable :: nper(:)
        call open_input(u, path)
        read (u, *) n0
        if (n0 /= na) call fatal(1, 'different number of atoms in "'//path//'" and the structure')
        read (u, *) rc
        allocate (nper(na))
        do a = 1, na
            read (u, *) nper(a)
            do k = 1, nper(a)
                do j = 1, 15
                    read (u, *)
                end do
            end do
        end do
        rewind (u)
        read (u, *) n0
        read (u, *) rc
        nt = sum(nper)
        allocate (a1(nt), a2(nt), a3(nt), lv2(3, nt), lv3(3, nt), m(3, 3, 3, nt))
        t = 0
        do a = 1, na
            read (u, *) ncount
            do k = 1, ncount
                t = t + 1
                a1(t) = a - 1
                read (u, *) idum
                read (u, *) a2(t)
                read (u, *) a3(t)
                a2(t) = a2(t) - 1
                a3(t) = a3(t) - 1
                read (u, *) v
                read (u, *) v
                lv2(:, t) = chop(matmul(latvec, v), sqtol)
                read (u, *) v
                lv3(:, t) = chop(matmul(latvec, v), sqtol)
                do i = 1, 3
                    do j = 1, 3
                        read (u, *) m(i, j, :, t)
                    end do
                end do
            end do
        end do
        close (u)
        m = m*ev_to_hartree/a_to_bohr**3
    end subroutine

    ! outfile.thermal_conductivity (main.f90:246-296); kappas in W/m/K, (3, 3).
    subroutine write_outfile(path, temperature, ksma, kcoll, koff, ktot)
        character(len=*), intent(in) :: path
        real(c_double), intent(in) :: temperature, ksma(3, 3), kcoll(3, 3), koff(3, 3), ktot(3, 3)
        integer :: u
        open (newunit=u, file=path, status='replace', action='write')
        write (u, '(A2,A5,15X,A)') '# ', 'Unit:', 'W/m/K'
        write (u, '(A2,A12,8X,E20.12)') '# ', 'Temperature:', temperature
        call block('# Single mode approximation', ksma)
        call block('# Collective contribution', kcoll)
        call block('# Off diagonal (coherence) contribution', koff)
        call block('# Total thermal conductivity', ktot)
        close (u)
    contains
        subroutine block(title, k)
            character(len=*), intent(in) :: title
            real(c_double), intent(in) :: k(3, 3)
            write (u, '(A)') title
            write (u, '(A1,6(1X,A24))') '#', 'kxx', 'kyy', 'kzz', 'kxy', 'kxz', 'kyz'
            write (u, '(1X,6(1X,E24.12))') k(1, 1), k(2, 2), k(3, 3), k(1, 2), k(1, 3), k(2, 3)
        end subroutine
    end subroutine

    ! ".T<temperature>" with the shortest decimal form of the temperature
    function temperature_suffix(t) result(s)
        real(c_double), intent(in) :: t
        character(len=:), allocatable :: s
        character(len=64) :: buf
        integer :: n
        if (t == anint(t) .and. abs(t) < 1.0e9_c_double) then
            write (buf, '(I0)') nint(t)
        else
            write (buf, '(F0.6)') t
            n = len_trim(buf)
            do while (buf(n:n) == '0')
                buf(n:n) = ' '
                n = n - 1
            end do
        end if
        s = '.T'//trim(buf)
    end function
end module tc_host_io

program tc_host
    use iso_c_binding
    use tc3_api
    use tc_host_io
    implicit none
    ! konstanter.f90: lo_kappa_au_to_SI
    real(c_double), parameter :: hartree_to_joule = 27.21138602_c_double*1.6021766208e-19_c_double
    real(c_double), parameter :: time_au_to_s = (6.582119514e-16_c_double*1.6021766208e-19_c_double)/hartree_to_joule
    real(c_double), parameter :: kappa_au_to_si = hartree_to_joule/time_au_to_s/0.52917721067e-10_c_double
    integer(c_int) :: qg(3), qg3(3), qg4(3), integrationtype, maxsteps, prec, nT, seed, mfppts, freqpts, dosint
    real(c_double) :: sigma, tol, max_mfp, dossigma
    character(len=3) :: unit
    character(len=:), allocatable :: elist
    real(c_double), allocatable :: temps(:)
    logical :: classical, noisotope, notr, offdiag, inputs_only, fourthorder, readiso, nothirdorder, nosym
    integer(c_int) :: device
    type(c_ptr) :: ctx
    integer :: na, npair, nfc3, iT, st
    real(c_double) :: latvec(3, 3), wmin, times(8)
    real(c_double), allocatable :: rfrac(:, :), mass(:), disorder(:), lv(:, :), m2(:, :, :), lv2(:, :), lv3(:, :), &
                                   m3(:, :, :, :)
    integer(c_int), allocatable :: species(:), p1(:), p2(:), t1(:), t2(:), t3(:), f1(:), f2(:), f3(:), f4(:)
    real(c_double), allocatable :: fl2(:, :), fl3(:, :), fl4(:, :), m4(:, :, :, :, :)
    integer :: nfc4
    real(c_double), allocatable, target :: ksma(:, :, :), kiter(:, :, :), kod(:, :, :)
    integer(c_int), allocatable, target :: nit(:)
    character(len=:), allocatable :: sfx

    call parse_options()
    call read_ucposcar('infile.ucposcar', na, latvec, rfrac, species, elist)
    allocate (mass(na), disorder(na))
    ! masses from infile.isotopes with --readiso, else natural isotopes (main.f90:98-104,
    ! type_crystalstructure.f90:650-652)
    if (readiso) then
        call read_isotopes('infile.isotopes', na, mass, disorder)
    else
        call natural_masses(na, species, elist, mass, disorder)
    end if
    call read_fc2('infile.forceconstant', na, latvec, npair, p1, p2, lv, m2)
    call read_fc3('infile.forceconstant_thirdorder', na, latvec, nfc3, t1, t2, t3, lv2, lv3, m3)

    call check(tc3_create(ctx, device), 'tc3_create')
    call check(tc3_set_structure(ctx, int(na, c_int), latvec, rfrac, species, mass), 'tc3_set_structure')
    call check(tc3_set_isotopes(ctx, disorder, merge(0_c_int, 1_c_int, noisotope)), 'tc3_set_isotopes')
    call check(tc3_set_thirdorder(ctx, merge(0_c_int, 1_c_int, nothirdorder)), 'tc3_set_thirdorder')
    ! options.f90:218: metres to Bohr (lo_m_to_Bohr)
    call check(tc3_set_boundary(ctx, merge(max_mfp/0.52917721067e-10_c_double, -1.0_c_double, max_mfp > 0)), &
               'tc3_set_boundary')
    call check(tc3_set_fc2(ctx, int(npair, c_int), p1, p2, lv, m2), 'tc3_set_fc2')
    call check(tc3_set_fc3(ctx, int(nfc3, c_int), t1, t2, t3, lv2, lv3, m3), 'tc3_set_fc3')
    if (fourthorder) then
        call read_fc4('infile.forceconstant_fourthorder', na, latvec, nfc4, f1, f2, f3, f4, fl2, fl3, fl4, m4)
        call check(tc3_set_fc4(ctx, int(nfc4, c_int), f1, f2, f3, f4, fl2, fl3, fl4, m4), 'tc3_set_fc4')
    end if
    call check(tc3_set_qmesh_sym(ctx, qg, merge(0_c_int, 1_c_int, notr), merge(0_c_int, 1_c_int, nosym)), &
               'tc3_set_qmesh_sym')
    call check(tc3_set_montecarlo(ctx, qg3, seed), 'tc3_set_montecarlo')
    call check(tc3_set_montecarlo4(ctx, qg4), 'tc3_set_montecarlo4')
    if (inputs_only) then
        call report_inputs()
    else
        call run_path()
    end if
    call tc3_destroy(ctx)

contains

    subroutine run_path()
        call check(tc3_harmonic(ctx, wmin), 'tc3_harmonic')
        call check(tc3_scatter(ctx, nT, temps, prec, integrationtype, sigma, merge(1_c_int, 0_c_int, classical)), &
                   'tc3_scatter')
        allocate (ksma(3, 3, nT), kiter(3, 3, nT), kod(3, 3, nT), nit(nT))
        kod = 0.0_c_double
        if (offdiag) then
            call check(tc3_kappa(ctx, maxsteps, tol, merge(1_c_int, 0_c_int, classical), c_loc(ksma), c_loc(kiter), &
                                 c_loc(kod), c_loc(nit)), 'tc3_kappa')
        else
            call check(tc3_kappa(ctx, maxsteps, tol, merge(1_c_int, 0_c_int, classical), c_loc(ksma), c_loc(kiter), &
                                 c_null_ptr, c_loc(nit)), 'tc3_kappa')
        end if
        do iT = 1, nT
            sfx = ''
            if (iT > 1) sfx = temperature_suffix(temps(iT))
            call write_outfile('outfile.thermal_conductivity'//sfx, temps(iT), ksma(:, :, iT)*kappa_au_to_si, &
                               (kiter(:, :, iT) - ksma(:, :, iT))*kappa_au_to_si, kod(:, :, iT)*kappa_au_to_si, &
                               (kiter(:, :, iT) + kod(:, :, iT))*kappa_au_to_si)
            call check(tc3_write_grid_hdf5(ctx, int(iT - 1, c_int), &
                                           'outfile.thermal_conductivity_grid'//sfx//'.hdf5'//c_null_char), &
                       'tc3_write_grid_hdf5')
            call check(tc3_write_cumulative_hdf5(ctx, int(iT - 1, c_int), &
                                                 'outfile.cumulative_thermal_conductivity'//sfx//'.hdf5'//c_null_char, &
                                                 mfppts, freqpts, dosint, sigma, dossigma, trim(unit)//c_null_char, &
                                                 elist//c_null_char), 'tc3_write_cumulative_hdf5')
        end do
        st = tc3_get_timings(ctx, times, 8_c_int)
        write (*, '(A,F0.3,A,F0.3,A,F0.3,A)', advance='no') 'harmonic ', times(1), ' s, scattering ', times(2), &
            ' s, kappa ', times(3), ' s (BTE iterations'
        do iT = 1, nT
            write (*, '(1X,I0)', advance='no') nit(iT)
        end do
        write (*, '(A)') ')'
    end subroutine

    ! sizes of the mesh and triplet sets, and checksums of what the readers produced
    subroutine report_inputs()
        integer(c_int) :: nsym, nfull, nirr, nrows
        integer(c_int64_t) :: ntrip
        call check(tc3_mesh_info(ctx, nsym, nfull, nirr, ntrip, nrows), 'tc3_mesh_info')
        write (*, '(A,5(1X,I0))') 'mesh', nsym, nfull, nirr, ntrip, nrows
        write (*, '(A,1X,I0,3(1X,ES24.16))') 'structure', na, sum(abs(latvec)), sum(rfrac), sum(mass)
        write (*, '(A,1X,ES24.16)') 'disorder', sum(disorder)
        write (*, '(A,1X,I0,3(1X,ES24.16))') 'fc2', npair, sum(abs(m2)), sum(m2(1, 2, :)), sum(abs(lv))
        write (*, '(A,1X,I0,4(1X,ES24.16))') 'fc3', nfc3, sum(abs(m3)), sum(m3(1, 2, 3, :)), sum(abs(lv2)), &
            sum(abs(lv3))
    end subroutine

    subroutine check(status, what)
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: what
        if (status /= 0) call fatal(int(status), what//': '//tc3_error_message(ctx))
    end subroutine

    subroutine parse_options()
        character(len=256) :: arg, val
        integer :: i, n, k
        real(c_double) :: tbuf(1024)
        qg = [26, 26, 26]
        qg3 = [-1, -1, -1]
        qg4 = [-1, -1, -1]
        seed = 0
        integrationtype = 2
        sigma = 1.0_c_double
        max_mfp = -1.0_c_double
        mfppts = 1000
        freqpts = 1000
        dosint = 3
        dossigma = 1.0_c_double
        unit = 'thz'
        tol = 1.0e-5_c_double
        maxsteps = 200
        prec = TC3_PREC_FP64
        classical = .false.
        noisotope = .false.
        notr = .false.
        readiso = .false.
        nothirdorder = .false.
        nosym = .false.
        offdiag = .true.
        inputs_only = .false.
        fourthorder = .false.
        device = 0
        nT = 1
        tbuf(1) = 300.0_c_double
        n = command_argument_count()
        i = 1
        do while (i <= n)
            call get_command_argument(i, arg)
            select case (trim(arg))
            case ('-qg', '--qpoint_grid')
                do k = 1, 3
                    call get_command_argument(i + k, val)
                    read (val, *) qg(k)
                end do
                i = i + 3
            case ('--temperature')
                nT = 0
                do while (i + 1 <= n)
                    call get_command_argument(i + 1, val)
                    if (val(1:1) == '-') exit
                    nT = nT + 1
                    read (val, *) tbuf(nT)
                    i = i + 1
                end do
                if (nT == 0) call fatal(5, '--temperature needs at least one value')
            case ('-qg3ph', '--qpoint_grid3ph')
                do k = 1, 3
                    call get_command_argument(i + k, val)
                    read (val, *) qg3(k)
                end do
                i = i + 3
            case ('-qg4ph', '--qpoint_grid4ph')
                do k = 1, 3
                    call get_command_argument(i + k, val)
                    read (val, *) qg4(k)
                end do
                i = i + 3
            case ('--seed')
                call get_command_argument(i + 1, val)
                read (val, *) seed
                i = i + 1
            case ('-it', '--integrationtype')
                call get_command_argument(i + 1, val)
                read (val, *) integrationtype
                i = i + 1
            case ('--sigma')
                call get_command_argument(i + 1, val)
                read (val, *) sigma
                i = i + 1
            case ('--max_mfp')
                call get_command_argument(i + 1, val)
                read (val, *) max_mfp
                i = i + 1
            case ('--mfppts')
                call get_command_argument(i + 1, val)
                read (val, *) mfppts
                i = i + 1
            case ('--freqpts')
                call get_command_argument(i + 1, val)
                read (val, *) freqpts
                i = i + 1
            case ('--dosintegrationtype')
                call get_command_argument(i + 1, val)
                read (val, *) dosint
                i = i + 1
            case ('--dossigma')
                call get_command_argument(i + 1, val)
                read (val, *) dossigma
                i = i + 1
            case ('--unit')
                call get_command_argument(i + 1, val)
                unit = trim(val)
                i = i + 1
            case ('--iterative_tolerance')
                call get_command_argument(i + 1, val)
                read (val, *) tol
                i = i + 1
            case ('--iterative_maxsteps')
                call get_command_argument(i + 1, val)
                read (val, *) maxsteps
                i = i + 1
            case ('--precision')
                call get_command_argument(i + 1, val)
                select case (trim(val))
                case ('fp64'); prec = TC3_PREC_FP64
                case ('fp32'); prec = TC3_PREC_FP32
                case ('bf16x3'); prec = TC3_PREC_BF16X3
                case default; call fatal(5, 'unknown precision "'//trim(val)//'"')
                end select
                i = i + 1
            case ('--classical'); classical = .true.
            case ('--noisotope'); noisotope = .true.
            case ('--notr'); notr = .true.
            case ('--readiso'); readiso = .true.
            case ('--nothirdorder'); nothirdorder = .true.
            case ('--nosym'); nosym = .true.
            case ('--no-offdiag'); offdiag = .false.
            case ('--fourthorder'); fourthorder = .true.
            case ('--inputs-only'); inputs_only = .true.
            case ('--device')
                call get_command_argument(i + 1, val)
                read (val, *) device
                i = i + 1
            case default
                call fatal(5, 'unknown option "'//trim(arg)//'"')
            end select
            i = i + 1
        end do
        temps = tbuf(1:nT)
        ! options.f90:220-223
        if (any(qg4 > 0) .and. .not. fourthorder) then
            write (*, *) 'You have to enable fourthorder to use a fourth order Monte-Carlo grid, stopping calculation.'
            stop
        end if
    end subroutine
end program tc_host
