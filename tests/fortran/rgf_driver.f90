! Test driver for the Fortran shim: an SCBA-style caller that solves every
! energy twice, once per energy through green_RGF_RS_gpu (the reference's
! argument list, as in the serial loop of src/green_rgf.f90:3174-3180) and once
! through the batched green_RGF_RS_energies_gpu, reading its inputs from in.bin
! and writing the results to out.bin (stream access, native byte order).
program rgf_driver
  use rgf_shim
  implicit none
  integer(4) :: nm, nx, nen, ie
  real(8) :: temp(2), mu(2)
  real(8), allocatable :: En(:), tr(:), tre(:), trb(:), treb(:)
  complex(8), allocatable :: Hii(:,:,:), H1i(:,:,:), sigL(:,:,:,:), sigG(:,:,:,:), sigR(:,:,:,:)
  complex(8), allocatable :: ldos(:,:,:,:), ndens(:,:,:,:), pdens(:,:,:,:), gri1(:,:,:,:), gli1(:,:,:,:), ggi1(:,:,:,:)
  complex(8), allocatable :: ldosb(:,:,:,:), ndensb(:,:,:,:), pdensb(:,:,:,:)
  real(8), allocatable :: cur(:,:,:,:), curb(:,:,:,:)
  open(unit=10, file='in.bin', access='stream', form='unformatted', status='old')
  read(10) nm, nx, nen
  allocate(En(nen), tr(nen), tre(nen), trb(nen), treb(nen))
  allocate(Hii(nm,nm,nx), H1i(nm,nm,nx))
  allocate(sigL(nm,nm,nx,nen), sigG(nm,nm,nx,nen), sigR(nm,nm,nx,nen))
  read(10) temp, mu, En, Hii, H1i, sigL, sigG, sigR
  close(10)
  allocate(ldos(nm,nm,nx,nen), ndens(nm,nm,nx,nen), pdens(nm,nm,nx,nen), cur(nm,nm,nx,nen))
  allocate(gri1(nm,nm,nx,nen), gli1(nm,nm,nx,nen), ggi1(nm,nm,nx,nen))
  allocate(ldosb(nm,nm,nx,nen), ndensb(nm,nm,nx,nen), pdensb(nm,nm,nx,nen), curb(nm,nm,nx,nen))
  gri1 = (0.0d0, 0.0d0); gli1 = (0.0d0, 0.0d0); ggi1 = (0.0d0, 0.0d0)
  do ie = 1, nen
    call green_RGF_RS_gpu(temp, nm, nx, En(ie), mu, Hii, H1i, sigL(:,:,:,ie), sigG(:,:,:,ie), sigR(:,:,:,ie), &
        ndens(:,:,:,ie), pdens(:,:,:,ie), ldos(:,:,:,ie), tr(ie), tre(ie), cur(:,:,:,ie), &
        gri1(:,:,:,ie), gli1(:,:,:,ie), ggi1(:,:,:,ie))
  end do
  call green_RGF_RS_energies_gpu(temp, nm, nx, nen, En, mu, Hii, H1i, sigL, sigG, sigR, &
      ndensb, pdensb, ldosb, trb, treb, curb)
  open(unit=11, file='out.bin', access='stream', form='unformatted', status='replace')
  write(11) ldos, ndens, pdens, cur, gri1, gli1, ggi1, tr, tre, ldosb, ndensb, pdensb, curb, trb, treb
  close(11)
  print '(A,I0,A)', 'rgf_driver: ', nen, ' energies done'
end program rgf_driver
