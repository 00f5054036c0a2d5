! Per-energy drop-in called from an OpenMP energy loop, as the reference's drivers
! call green_RGF_RS from every thread (src/green_rgf.f90:1301-1318): each thread
! calls green_RGF_RS_gpu for its energies.  Reads in.bin (nm, nx, nen, temp, mu,
! En, Hii, H1i; stream access, native byte order), solves the ballistic device
! (zero scattering self-energies), prints the wall time of the loop and writes
! tr, tre and ldos to out.bin.  The thread count is OMP_NUM_THREADS.
program rgf_omp_driver
  use rgf_shim
  use omp_lib
  implicit none
  integer(4) :: nm, nx, nen, ie
  real(8) :: temp(2), mu(2), t0, t1
  real(8), allocatable :: En(:), tr(:), tre(:)
  complex(8), allocatable :: Hii(:,:,:), H1i(:,:,:), sig0(:,:,:)
  complex(8), allocatable :: ldos(:,:,:,:), ndens(:,:,:,:), pdens(:,:,:,:)
  real(8), allocatable :: cur(:,:,:,:)
  open(unit=10, file='in.bin', access='stream', form='unformatted', status='old')
  read(10) nm, nx, nen
  allocate(En(nen), tr(nen), tre(nen))
  allocate(Hii(nm,nm,nx), H1i(nm,nm,nx), sig0(nm,nm,nx))
  read(10) temp, mu, En, Hii, H1i
  close(10)
  sig0 = (0.0d0, 0.0d0)
  allocate(ldos(nm,nm,nx,nen), ndens(nm,nm,nx,nen), pdens(nm,nm,nx,nen), cur(nm,nm,nx,nen))
  ! one call before the timed loop: library load and device workspace
  call green_RGF_RS_gpu(temp, nm, nx, En(1), mu, Hii, H1i, sig0, sig0, sig0, &
      ndens(:,:,:,1), pdens(:,:,:,1), ldos(:,:,:,1), tr(1), tre(1), cur(:,:,:,1))
  t0 = omp_get_wtime()
  !$omp parallel do schedule(static, 1)
  do ie = 1, nen
    call green_RGF_RS_gpu(temp, nm, nx, En(ie), mu, Hii, H1i, sig0, sig0, sig0, &
        ndens(:,:,:,ie), pdens(:,:,:,ie), ldos(:,:,:,ie), tr(ie), tre(ie), cur(:,:,:,ie))
  end do
  !$omp end parallel do
  t1 = omp_get_wtime()
  open(unit=11, file='out.bin', access='stream', form='unformatted', status='replace')
  write(11) tr, tre, ldos
  close(11)
  print '(A,I0,A,I0,A,F12.6)', 'rgf_omp_driver: threads ', omp_get_max_threads(), ' energies ', nen, &
      ' seconds ', t1 - t0
end program rgf_omp_driver
