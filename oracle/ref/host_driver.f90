! TEST INFRASTRUCTURE ONLY.
!
! The drop-in check: Fortran calling IntertwineCPP exactly as the
! reference's GPU_Intertwine does (GPUIterate.f90:130-180: block packed into
! qbase/qcoef with the 1e-14 coefficient noise, the HopList arrays, the
! LOGICAL update, every argument by reference, IntertwineCPP an implicit
! external), linked against libmlcp_hip.so, whose intertwinecpp_ it
! resolves to.  GPU_Intertwine itself is not called: it frees each state
! (FlushCPvec, GPUIterate.f90:145) before checking SIZE(Q(j)%coef) against
! the rank (:153), which is undefined; compiled with amdflang it reads 0 and
! aborts with "bad rank".  The packing here checks the ranks first.
!
! Usage: host_driver <input.bin> <output.bin>
! Input: as ref_driver (uniform ranks); npow, nals, update, ishift, eshift
! from the header.  Output: int32 status 0, then real64 coef (rk*nbloc),
! base (nD, rk*nbloc), eigv (nbloc).

      program host_driver
      use SEPDREPN
      use HAMILSETUP
      use INPUTFIELDS
      implicit none
      character(len=512) :: fin,fout
      integer :: hdr(12),mode,ndof,nbloc,T,nops,msize,npow,nals,upd
      integer :: ishift,nextra,nD,ntot,s,rst
      real*8  :: eshift
      integer, allocatable :: nbas(:),ranks(:),opid(:),mptr(:)
      real*8, allocatable  :: hcoef(:),mat(:),coef(:),base(:,:),eigv(:)
      TYPE (HopList) :: HL
      TYPE (CPpar) :: cpp
      TYPE (CPvec), allocatable :: Q(:)
      real*8, allocatable :: hlcoef(:),hlmat(:),eig(:),qcoef(:),qbase(:,:)
      real*8 :: esft
      real*8, parameter :: noise=1.d-14
      integer :: rk

      call get_command_argument(1,fin)
      call get_command_argument(2,fout)
      open(10,file=trim(fin),access='stream',form='unformatted',&
           status='old')
      read(10) hdr
      IF (hdr(1).ne.1296843600) STOP 2
      mode=hdr(2); ndof=hdr(3); nbloc=hdr(4); T=hdr(5); nops=hdr(6)
      msize=hdr(7); npow=hdr(8); nals=hdr(9); upd=hdr(10)
      ishift=hdr(11); nextra=hdr(12)
      read(10) eshift
      allocate(nbas(ndof),ranks(nbloc),opid(T*ndof),mptr(nops))
      allocate(hcoef(T),mat(msize))
      read(10) nbas,ranks,opid,mptr,hcoef,mat
      nD=SUM(nbas)
      ntot=SUM(ranks)
      allocate(coef(ntot),base(nD,ntot),eigv(nbloc))
      read(10) coef,base,eigv
      close(10)

      allocate(HL%opid(T*ndof),HL%mptr(nops),HL%coef(T),HL%mat(msize))
      HL%opid=opid
      HL%mptr=mptr
      HL%coef=hcoef
      HL%mat=mat

      cpp%npow=npow
      cpp%psinals=nals
      cpp%psirank=ranks(1)
      cpp%ncpu=1
      cpp%update=(upd.ne.0)

      allocate(Q(nbloc))
      rst=0
      DO s=1,nbloc
         call NewCPvec(Q(s),nbas,ranks(s))
         Q(s)%coef(:)=coef(rst+1:rst+ranks(s))
         Q(s)%base(:,:)=base(:,rst+1:rst+ranks(s))
         rst=rst+ranks(s)
      ENDDO

      rk=cpp%psirank
      DO s=1,nbloc
         IF (SIZE(Q(s)%coef).ne.rk) STOP 3
      ENDDO
      allocate(hlcoef(SIZE(HL%coef)),hlmat(SIZE(HL%mat)),eig(nbloc))
      allocate(qcoef(rk*nbloc),qbase(nD,rk*nbloc))
      hlcoef(:)=HL%coef(:)
      hlmat(:)=HL%mat(:)
      eig(:)=eigv(:)
      rst=0
      DO s=1,nbloc
         qcoef(rst+1:rst+rk)=Q(s)%coef(:)+noise
         qbase(:,rst+1:rst+rk)=Q(s)%base(:,:)
         rst=rst+rk
      ENDDO
      esft=eshift
      call IntertwineCPP(qbase,qcoef,nbas,&
                         rk,SIZE(nbas),nbloc,&
                         HL%opid,HL%mptr,hlcoef,hlmat,&
                         SIZE(HL%opid),SIZE(HL%mptr),SIZE(hlmat),&
                         eig,cpp%npow,cpp%ncpu,cpp%psinals,cpp%update,&
                         ishift,esft)
      eigv(:)=eig(:)
      rst=0
      DO s=1,nbloc
         Q(s)%coef(:)=qcoef(rst+1:rst+rk)
         Q(s)%base(:,:)=qbase(:,rst+1:rst+rk)
         rst=rst+rk
      ENDDO

      open(11,file=trim(fout),access='stream',form='unformatted',&
           status='replace')
      write(11) 0
      DO s=1,nbloc
         write(11) Q(s)%coef
      ENDDO
      DO s=1,nbloc
         write(11) Q(s)%base
      ENDDO
      write(11) eigv
      close(11)

      end program host_driver
