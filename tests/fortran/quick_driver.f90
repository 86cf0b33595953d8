! A Fortran caller of libquickfock through QUICK's own GPU entry points, with
! the implicit interfaces, by-reference arguments, LOGICAL flags and
! column-major arrays of QUICK's Fortran (main.f90:99-204,
! quick_scf_operator_module.f90:89-126, gradient.fh:575-584, 867).
!
!   quick_driver <input file> <output file>
!
! The input file (written by tests/test_gpu_fortran.py) holds, as plain
! numbers: natom nbasis nshell jbasis nelec; xyz(3,natom) in bohr; the
! nuclear charges; per shell kstart katom kprim ksumtype qstart qfinal
! (1-based); gccoeff(14,nbasis) and gcexpo(14,nbasis); ycutoff(nshell,nshell)
! and cutprim(jbasis,jbasis); the density dense(nbasis,nbasis) and its
! cutmatrix(nshell,nshell); integralCutoff, primLimit, gradCutoff.  The
! output file receives o = 0.25 + G(dense) (nbasis*nbasis numbers, column
! order) and then 0.125 + the two-electron gradient (3*natom numbers).
program quick_driver
   implicit none
   integer, parameter :: MAXPRIM = 14
   integer :: natom, nbasis, nshell, jbasis, nelec, ierr, i, j, s
   integer :: imult, molchg, method
   double precision, allocatable :: xyz(:,:), chg(:), gccoeff(:,:), gcexpo(:,:), ycut(:,:), cutprim(:,:)
   double precision, allocatable :: dense(:,:), cutmat(:,:), o(:,:), grad(:), dummyd(:)
   integer, allocatable :: kstart(:), katom(:), kprim(:), ksumtype(:), qstart(:), qfinal(:), dummyi(:), iatomtype(:)
   double precision :: icut, plim, gcut, hyb, dmcut, corecut
   logical :: is_oshell, deltao
   character(len=512) :: fin, fout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (unit=10, file=trim(fin), status='old', action='read')
   read (10, *) natom, nbasis, nshell, jbasis, nelec
   allocate (xyz(3, natom), chg(natom), iatomtype(natom))
   allocate (kstart(nshell), katom(nshell), kprim(nshell), ksumtype(nshell), qstart(nshell), qfinal(nshell))
   allocate (gccoeff(MAXPRIM, nbasis), gcexpo(MAXPRIM, nbasis), ycut(nshell, nshell), cutprim(jbasis, jbasis))
   allocate (dense(nbasis, nbasis), cutmat(nshell, nshell), o(nbasis, nbasis), grad(3*natom))
   allocate (dummyi(4*max(nbasis, nshell) + 8), dummyd(4*max(nbasis, 1)))
   read (10, *) xyz
   read (10, *) chg
   do s = 1, nshell
      read (10, *) kstart(s), katom(s), kprim(s), ksumtype(s), qstart(s), qfinal(s)
   end do
   read (10, *) gccoeff
   read (10, *) gcexpo
   read (10, *) ycut
   read (10, *) cutprim
   read (10, *) dense
   read (10, *) cutmat
   read (10, *) icut, plim, gcut
   close (10)
   dummyi = 0
   dummyd = 0.0d0
   iatomtype = 1

   call gpu_new(ierr)
   if (ierr /= 0) stop 2
   call gpu_set_device(0, ierr)
   call gpu_init_device(ierr)
   if (ierr /= 0) stop 3
   method = 0
   is_oshell = .false.
   hyb = 1.0d0
   call gpu_upload_method(method, is_oshell, hyb)
   imult = 1
   molchg = 0
   call gpu_setup(natom, nbasis, nelec, imult, molchg, iatomtype)
   call gpu_upload_xyz(xyz)
   call gpu_upload_atom_and_chg(iatomtype, chg)
   call gpu_upload_basis(nshell, jbasis, nshell, jbasis, MAXPRIM, dummyi, dummyi, dummyd, dummyd, &
                         dummyi, dummyi, dummyi, dummyi, dummyi, kstart, katom, dummyi, kprim, dummyi, &
                         ksumtype, dummyi, qstart, qfinal, dummyi, dummyi, gccoeff, dummyd, gcexpo, dummyi)
   call gpu_upload_cutoff_matrix(ycut, cutprim)

   ! one SCF-operator call: o += G(dense)
   o = 0.25d0
   call gpu_upload_calculated(o, dummyd, dummyd, dense)
   dmcut = 1.0d-10
   corecut = 1.0d-12
   call gpu_upload_cutoff(cutmat, icut, plim, dmcut, corecut)
   deltao = .false.
   call gpu_get_cshell_eri(deltao, o)

   ! the gradient driver's sequence: density, cutoffs, gradCutoff, ERI gradient
   grad = 0.125d0
   call gpu_upload_density_matrix(dense)
   call gpu_upload_cutoff(cutmat, icut, plim, dmcut, corecut)
   call gpu_upload_grad(gcut)
   call gpu_get_cshell_eri_grad(grad)
   call gpu_delete(ierr)

   open (unit=11, file=trim(fout), status='replace', action='write')
   do j = 1, nbasis
      do i = 1, nbasis
         write (11, '(ES26.17E3)') o(i, j)
      end do
   end do
   do i = 1, 3*natom
      write (11, '(ES26.17E3)') grad(i)
   end do
   close (11)
end program quick_driver
