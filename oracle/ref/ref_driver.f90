! TEST INFRASTRUCTURE ONLY.
!
! Driver that runs the compiled reference MLCP FP64 CPU routines (lowmem=3)
! on inputs captured at the intertwinecpp_ boundary, so that the oracle and
! the HIP library can be checked against the reference itself.
!
! Usage: ref_driver <input.bin> <output.bin>
!
! Input (stream, little endian): int32 header(12) = magic 1296843600, mode,
!   ndof, nbloc, T, nops, msize, npow, nals, update, ishift, nextra;
!   real64 eshift; int32 nbas(ndof), ranks(nbloc), opid(T*ndof), mptr(nops);
!   real64 hcoef(T), mat(msize), coef(sum ranks), base(nD*sum ranks),
!   eigv(nbloc), extra(nextra).
! Modes: 1 ALS_POW3 per state; 2 PRODHV_ALS3 per state; 3 ALS_ORTHO_3;
!   4 GetQHQ3 + GetOverlaps; 5 one Iterate cycle (styp=-1, update as given);
!   6 UpdateVecs2 with extra = U(nbloc,nbloc); 7 reduc_ALS(G=state 2,
!   F=state 1); 8 RayleighQuotient2 per state; 9 SolveGenEigval with
!   extra = [QHQ, S]; 10 UpdateOperMat of every operator of the term list on
!   its mode, in operator order (each output is the packed nbloc x nbloc
!   matrix); 11 SavePsiFile of the block to <output>.psi with isavi = npow,
!   extra = [bounds(2), delta(nbloc)]; 12 ReadPsiFile of <input>.psi into
!   the block, output [isavi, success, bounds(2), eigv, delta] and the block
!   read; 13 ReadFFHamiltonian of pes/f<k><id>.dat in the working directory
!   (id = 5 character codes in extra, ncp = npow, divide = update), output
!   per order k: [n_k, qns(n_k,k), coef(n_k)]; 14 GetHopList of a
!   Hamiltonian given as opindx (in opid, signed), opcoef (hcoef), and the
!   operator matrices iops(ndof), eops(ndof), cops(npow), pops(nals) (modes
!   in mptr, packed triangles concatenated in mat), output
!   [size(opid), size(mptr), size(mat), opid, mptr, mat, coef].
! Modes 13 and 14 read no term list: the opid/mptr/mat records carry their
! own inputs as described.
! Output: int32 status, then real64 arrays depending on the mode.
!
! The term list is turned into a TYPE(Hamiltonian) whose every operator is a
! "primitive" operator matrix; an identity stored as a dense packed triangle
! gives bit-identical HVBaseProd results to the reference's opindx=0 path.

      program ref_driver
      use SEPDREPN
      use HAMILSETUP
      use ALSPOW
      use ALSUTILS
      use BLOCKUTILS
      use REDUCTION
      use LINALG
      use UPDATER
      use RESTART
      use FFPES
      use MODEH
      implicit none
      character(len=512) :: fin,fout
      integer :: hdr(12),mode,ndof,nbloc,T,nops,msize,npow,nals,upd
      integer :: ishift,nextra,nD,ntot,i,j,s,op,rst,nq
      real*8  :: eshift
      integer, allocatable :: nbas(:),ranks(:),opid(:),mptr(:),opmode(:)
      real*8, allocatable  :: hcoef(:),mat(:),coef(:),base(:,:),eigv(:)
      real*8, allocatable  :: extra(:),QHQ(:,:),Smat(:,:),rq(:)
      real*8  :: bnds(2)
      real*8, allocatable  :: dlt(:)
      character(len=64) :: fpsi
      character(len=5) :: ffid
      logical :: ok
      TYPE (Configs), allocatable :: W(:)
      TYPE (HopList) :: HL
      integer :: nq2,ncop,npop
      TYPE (Hamiltonian) :: H
      TYPE (CPvec), allocatable :: Q(:),HQ(:)

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
      allocate(hcoef(T),mat(msize),opmode(nops))
      read(10) nbas,ranks,opid,mptr,hcoef,mat
      nD=SUM(nbas)
      ntot=SUM(ranks)
      allocate(coef(ntot),base(nD,ntot),eigv(nbloc),extra(max(nextra,1)))
      read(10) coef,base,eigv
      IF (nextra.gt.0) read(10) extra(1:nextra)
      close(10)

      IF (mode.eq.13 .or. mode.eq.14) THEN
         open(11,file=trim(fout),access='stream',form='unformatted',&
              status='replace')
         IF (mode.eq.13) THEN
            DO i=1,5
               ffid(i:i)=achar(nint(extra(i)))
            ENDDO
            call ReadFFHamiltonian(W,ffid,npow,upd.ne.0)
            write(11) 0
            DO i=1,npow
               write(11) dble(SIZE(W(i)%coef)),dble(W(i)%qns),W(i)%coef
            ENDDO
         ELSE
            ncop=npow
            npop=nals
            allocate(H%opcoef(T),H%opindx(T,ndof),H%iops(ndof),&
                     H%eops(ndof),H%cops(ncop),H%pops(npop))
            H%opcoef(:)=hcoef(:)
            DO i=1,T
               DO j=1,ndof
                  H%opindx(i,j)=opid((i-1)*ndof+j)
               ENDDO
            ENDDO
            rst=1
            DO op=1,nops
               nq=nbas(mptr(op))*(nbas(mptr(op))+1)/2
               IF (op.le.ndof) THEN
                  allocate(H%iops(op)%mat(nq))
                  H%iops(op)%mat(:)=mat(rst:rst+nq-1)
                  H%iops(op)%dof=mptr(op)
               ELSEIF (op.le.2*ndof) THEN
                  allocate(H%eops(op-ndof)%mat(nq))
                  H%eops(op-ndof)%mat(:)=mat(rst:rst+nq-1)
                  H%eops(op-ndof)%dof=mptr(op)
               ELSEIF (op.le.2*ndof+ncop) THEN
                  allocate(H%cops(op-2*ndof)%mat(nq))
                  H%cops(op-2*ndof)%mat(:)=mat(rst:rst+nq-1)
                  H%cops(op-2*ndof)%dof=mptr(op)
               ELSE
                  allocate(H%pops(op-2*ndof-ncop)%mat(nq))
                  H%pops(op-2*ndof-ncop)%mat(:)=mat(rst:rst+nq-1)
                  H%pops(op-2*ndof-ncop)%dof=mptr(op)
               ENDIF
               rst=rst+nq
            ENDDO
            call GetHopList(H,HL,ndof,.TRUE.)
            write(11) 0
            write(11) dble(SIZE(HL%opid)),dble(SIZE(HL%mptr)),&
                      dble(SIZE(HL%mat)),dble(HL%opid),dble(HL%mptr),&
                      HL%mat,HL%coef
         ENDIF
         close(11)
         STOP
      ENDIF

!     Term list -> TYPE(Hamiltonian)
      opmode=0
      DO i=1,T
         DO j=1,ndof
            opmode(opid((i-1)*ndof+j))=j
         ENDDO
      ENDDO
      allocate(H%opcoef(T),H%opindx(T,ndof),H%pops(nops),H%eops(0),&
               H%cops(0))
      H%opcoef(:)=hcoef(:)
      DO i=1,T
         DO j=1,ndof
            H%opindx(i,j)=opid((i-1)*ndof+j)
         ENDDO
      ENDDO
      DO op=1,nops
         IF (opmode(op).eq.0) THEN
            allocate(H%pops(op)%mat(1))
            H%pops(op)%mat=0.d0
         ELSE
            nq=nbas(opmode(op))*(nbas(opmode(op))+1)/2
            allocate(H%pops(op)%mat(nq))
            H%pops(op)%mat(:)=mat(mptr(op):mptr(op)+nq-1)
         ENDIF
      ENDDO

!     Block -> TYPE(CPvec)
      allocate(Q(nbloc))
      rst=0
      DO s=1,nbloc
         call NewCPvec(Q(s),nbas,ranks(s))
         Q(s)%coef(:)=coef(rst+1:rst+ranks(s))
         Q(s)%base(:,:)=base(:,rst+1:rst+ranks(s))
         rst=rst+ranks(s)
      ENDDO

      open(11,file=trim(fout),access='stream',form='unformatted',&
           status='replace')
      SELECT CASE (mode)
      CASE (1)
!        states in parallel, as Iterate does (Solver.f90:509-521); a
!        comment unless built with -fopenmp (ref_driver_omp)
!$omp parallel do schedule(dynamic)
         DO s=1,nbloc
            call ALS_POW3(H,Q(s),npow,ishift,eshift)
         ENDDO
!$omp end parallel do
         call putblock()
      CASE (2)
         allocate(HQ(nbloc))
!$omp parallel do schedule(dynamic)
         DO s=1,nbloc
            call PRODHV_ALS3(Q(s),HQ(s),H,ishift,eshift,nals)
         ENDDO
!$omp end parallel do
         write(11) 0
         DO s=1,nbloc
            write(11) HQ(s)%coef
         ENDDO
         DO s=1,nbloc
            write(11) HQ(s)%base
         ENDDO
      CASE (3)
         call ALS_ORTHO_3(Q,nals)
         call putblock()
      CASE (4)
         allocate(QHQ(nbloc,nbloc),Smat(nbloc,nbloc))
         call GetQHQ3(Q,H,QHQ,nals,3)
         call GetOverlaps(Q,Smat)
         write(11) 0
         write(11) QHQ,Smat
      CASE (5)
         DO s=1,nbloc
            call ALS_POW_alg(H,Q(s),npow,1,eshift,3)
         ENDDO
         call ALS_ORTHO_alg(Q,nals,3)
         IF (upd.ne.0) call Diagonalize(Q,H,eigv,.TRUE.,nals,3)
         call putblock()
         write(11) eigv
      CASE (6)
         allocate(QHQ(nbloc,nbloc))
         QHQ=reshape(extra(1:nbloc*nbloc),(/nbloc,nbloc/))
         call UpdateVecs2(Q,QHQ,nals,3)
         call putblock()
      CASE (7)
         call reduc_ALS(Q(2),Q(1),nals)
         write(11) 0
         write(11) Q(1)%coef,Q(1)%base
      CASE (8)
         allocate(rq(nbloc))
         DO s=1,nbloc
            rq(s)=RayleighQuotient2(Q(s),H)
         ENDDO
         write(11) 0
         write(11) rq
      CASE (9)
         allocate(QHQ(nbloc,nbloc),Smat(nbloc,nbloc))
         QHQ=reshape(extra(1:nbloc*nbloc),(/nbloc,nbloc/))
         Smat=reshape(extra(nbloc*nbloc+1:2*nbloc*nbloc),(/nbloc,nbloc/))
         call SolveGenEigval(eigv,Smat,QHQ,'V')
         write(11) 0
         write(11) eigv,QHQ
      CASE (10)
         write(11) 0
         DO op=1,nops
            IF (opmode(op).gt.0) THEN
               call UpdateOperMat(opmode(op),Q,H%pops(op))
               write(11) H%pops(op)%mat
            ENDIF
         ENDDO
      CASE (11)
         bnds=extra(1:2)
         allocate(dlt(nbloc))
         dlt=extra(3:2+nbloc)
         fpsi=trim(fout)//'.psi'
         call SavePsiFile(npow,bnds,eigv,dlt,Q,fpsi)
         write(11) 0
      CASE (12)
         allocate(dlt(nbloc))
         bnds=0.d0
         dlt=0.d0
         i=-1
         fpsi=trim(fin)//'.psi'
         call ReadPsiFile(i,bnds,eigv,dlt,Q,fpsi,ok)
         write(11) 0
         write(11) dble(i),merge(1.d0,0.d0,ok),bnds,eigv,dlt
         DO s=1,nbloc
            write(11) Q(s)%coef
         ENDDO
         DO s=1,nbloc
            write(11) Q(s)%base
         ENDDO
      CASE DEFAULT
         write(11) 1
      END SELECT
      close(11)

      contains

      subroutine putblock()
      integer :: ss
      write(11) 0
      DO ss=1,nbloc
         write(11) Q(ss)%coef
      ENDDO
      DO ss=1,nbloc
         write(11) Q(ss)%base
      ENDDO
      end subroutine putblock

      end program ref_driver
