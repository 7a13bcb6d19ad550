c     ref_driver.f -- bind(C) entry points that fill the reference's
c     COMMON blocks and call its own routines (search2, setpar, setparv,
c     setham, sethamv, diag, perturb, fermi, bande) unchanged.
c
c     TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/libnrlref.so
c     together with the reference's sources compiled where they lie
c     (NRLtb/ex1, whose P1/P2/P3 give matom=20, mbas=9, mpress=1).
c     Used to write golden vectors and to pin the C restatement.
c
c     Layouts seen from C (all arrays by pointer):
c       lat(3,3)   lat(c,v) = component c of lattice vector v
c       pos(3,nat) Cartesian positions (bohr); pq(3,nat) phase origins
c       packed H/S: complex*16 upper packed, nv*(nv+1)/2 entries
c
      subroutine ref_setup(nat,lat,pos,pq,kk,nd,nkinds,kb,val,
     $     dnnin,scrin,par,npar,nlt,jsov,npout,dmin,ierr)
     $     bind(C,name='ref_setup')
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer nat,nkinds,npar,nlt,jsov,npout,ierr
      integer kk(nat),nd(3),kb(nkinds)
      real*8 lat(3,3),pos(3,nat),pq(3,nat),val(nkinds)
      real*8 dnnin(*),scrin(*),par(npar),dmin
      logical master,multiprc
      common /MPIcomm/ MPISIZE,MPIRANK,master,multiprc
      common /codes/posn(matom,3),kkind(matom),natoms
      common /latt1/plv(3,3)
      common /parinfo/ valence(mkind),kinds,kbas(mkind)
      common /cutcom/ dnnsq(mint),dnn(mint),
     $     screenl(mint),scrninv(mint)
      integer nkdarry(6)
      common/disply/ nkdarry
      common /kstuff/ ksk(matom)
      common /types/ npkind(mptype)
      common /fazon/posq(matom,3)
      common /struc/ nv
      logical jsover
      common /lia/ npts,jsover
      common /params/ param(npard,mspin)
      logical realov,harrison
      common /parcom/ nltype,realov,harrison
      ierr = 0
      if(nat.gt.matom.or.nkinds.gt.mkind.or.npar.gt.npard) then
         ierr = 10
         return
      end if
      MPISIZE = 1
      MPIRANK = 0
      master = .true.
      multiprc = .false.
      npkind(1) = 4
      npkind(2) = 10
      npkind(3) = 14
      natoms = nat
      kinds = nkinds
      do i = 1,3
         do j = 1,3
            plv(j,i) = lat(i,j)
         end do
      end do
      do i = 1,nkinds
         kbas(i) = kb(i)
         valence(i) = val(i)
      end do
      do i = 1,nkinds*(nkinds+1)/2
         dnn(i) = dnnin(i)
         screenl(i) = scrin(i)
         dnnsq(i) = dnn(i)*dnn(i)
         scrninv(i) = 1d0/screenl(i)
      end do
      nv = 0
      do ia = 1,nat
         kkind(ia) = kk(ia)
         do j = 1,3
            posn(ia,j) = pos(j,ia)
            posq(ia,j) = pq(j,ia)
         end do
         ksk(ia) = nv
         nv = nv + kbas(kkind(ia))
      end do
      do i = 1,3
         nkdarry(i) = nd(i)
         nkdarry(i+3) = 2*nd(i)+1
      end do
      do i = 1,npard
         param(i,1) = 0d0
      end do
      do i = 1,npar
         param(i,1) = par(i)
      end do
      nltype = nlt
      realov = (nlt.eq.1).or.(nlt.eq.3)
      harrison = (nlt.eq.2).or.(nlt.eq.3)
      jsover = jsov.ne.0
      call search2(npout,dmin,ierr)
      return
      end
c
      subroutine ref_get_pairs(tto,dlvo,disto,scro,dscro,jko,jao)
     $     bind(C,name='ref_get_pairs')
      implicit real*8 (a-h,o-z)
      include 'P1'
      real*8 tto(3,*),dlvo(3,*),disto(*),scro(*),dscro(*)
      integer jko(2,*),jao(2,*)
      common /pairs/tt_list(3,mpair),
     $     dlv_list(3,mpair),dist_list(mpair),
     $     screen_list(mpair),
     $     jkind_list(mpair,2),
     $     jatm_list(mpair,2),
     $     npairs
      common /neigh/ dscrn(mpair)
      do ip = 1,npairs
         do j = 1,3
            tto(j,ip) = tt_list(j,ip)
            dlvo(j,ip) = dlv_list(j,ip)
         end do
         disto(ip) = dist_list(ip)
         scro(ip) = screen_list(ip)
         dscro(ip) = dscrn(ip)
         jko(1,ip) = jkind_list(ip,1)
         jko(2,ip) = jkind_list(ip,2)
         jao(1,ip) = jatm_list(ip,1)
         jao(2,ip) = jatm_list(ip,2)
      end do
      return
      end
c
c     Per-structure parameters; results kept here for the per-k calls.
c
      subroutine ref_setpar(thisvol,dpar,hpo,opo,dparo,dhpo,dopo,
     $     deno,ierr) bind(C,name='ref_setpar')
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer ierr
      real*8 dpar(4,*),hpo(14,*),opo(14,*),dparo(4,*),dhpo(14,*)
      real*8 dopo(14,*),deno(mkind,*)
      logical flipos(matom)
      do i = 1,matom
         flipos(i) = .false.
      end do
      call ref_setpar_js(1,flipos,thisvol,dpar,hpo,opo,dparo,dhpo,
     $     dopo,deno,ierr)
      return
      end
c
c     Spin channel jspin of a spin-polarised set: param(:,2) from
c     pdn(npar), atom ia flipped where iflip(ia) is non-zero.
c
      subroutine ref_setpar_spin(jspin,thisvol,pdn,npar,iflip,dpar,
     $     hpo,opo,dparo,dhpo,dopo,deno,ierr)
     $     bind(C,name='ref_setpar_spin')
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer jspin,npar,ierr,iflip(*)
      real*8 pdn(npar)
      real*8 dpar(4,*),hpo(14,*),opo(14,*),dparo(4,*),dhpo(14,*)
      real*8 dopo(14,*),deno(mkind,*)
      common /codes/posn(matom,3),kkind(matom),natoms
      common /params/ param(npard,mspin)
      logical flipos(matom)
      do i = 1,npard
         param(i,2) = 0d0
      end do
      do i = 1,npar
         param(i,2) = pdn(i)
      end do
      do i = 1,matom
         flipos(i) = .false.
      end do
      do i = 1,natoms
         flipos(i) = iflip(i).ne.0
      end do
      call ref_setpar_js(jspin,flipos,thisvol,dpar,hpo,opo,dparo,dhpo,
     $     dopo,deno,ierr)
      return
      end
c
c     Non-collinear set-up (lnoncol): param(:,2) from pdn(npar); returns
c     paros, spinos, dparos, dspnos, both spins' densities and channel 1's
c     bond and overlap integrals.
c
      subroutine ref_setpar_nc(thisvol,pdn,npar,dpar,spo,dparo,dspo,
     $     hpo,opo,dhpo,dopo,deno,ierr) bind(C,name='ref_setpar_nc')
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer npar,ierr
      real*8 pdn(npar),dpar(4,*),spo(4,*),dparo(4,*),dspo(4,*)
      real*8 hpo(14,*),opo(14,*),dhpo(14,*),dopo(14,*)
      real*8 deno(mkind,matom,2)
      common /codes/posn(matom,3),kkind(matom),natoms
      common /pairs/tt_list(3,mpair),
     $     dlv_list(3,mpair),dist_list(mpair),
     $     screen_list(mpair),
     $     jkind_list(mpair,2),
     $     jatm_list(mpair,2),
     $     npairs
      logical jsover
      common /lia/ npts,jsover
      common /params/ param(npard,mspin)
      common /refstate/ paros(4,matom),hop(mpkind,mpair),
     $     over(mpkind,mpair),dparos(4,matom),dhop(mpkind,mpair),
     $     dover(mpkind,mpair),den(mkind,matom,mspin)
      real*8 spinos(4,matom),dspnos(4,matom)
      logical flipos(matom)
      do i = 1,npard
         param(i,2) = 0d0
      end do
      do i = 1,npar
         param(i,2) = pdn(i)
      end do
      do i = 1,matom
         flipos(i) = .false.
         do k = 1,4
            spinos(k,i) = 0d0
            dspnos(k,i) = 0d0
         end do
      end do
      do ip = 1,npairs
         do j = 1,mpkind
            hop(j,ip) = 0d0
            over(j,ip) = 0d0
            dhop(j,ip) = 0d0
            dover(j,ip) = 0d0
         end do
      end do
      call setpar(1,param,jsover,den,flipos,.true.,
     $     paros,spinos,hop,over,.false.,ierr)
      if(ierr.ne.0) return
      call setparv(1,param,thisvol,jsover,den,flipos,.true.,
     $     dparos,dspnos,dhop,dover,.false.,ierr)
      do ia = 1,natoms
         do k = 1,4
            dpar(k,ia) = paros(k,ia)
            spo(k,ia) = spinos(k,ia)
            dparo(k,ia) = dparos(k,ia)
            dspo(k,ia) = dspnos(k,ia)
         end do
         do k = 1,mkind
            deno(k,ia,1) = den(k,ia,1)
            deno(k,ia,2) = den(k,ia,2)
         end do
      end do
      do ip = 1,npairs
         do j = 1,14
            hpo(j,ip) = hop(j,ip)
            opo(j,ip) = over(j,ip)
            dhpo(j,ip) = dhop(j,ip)
            dopo(j,ip) = dover(j,ip)
         end do
      end do
      return
      end
c
      subroutine ref_setpar_js(jspin,flipos,thisvol,dpar,hpo,opo,dparo,
     $     dhpo,dopo,deno,ierr)
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer jspin,ierr
      logical flipos(matom)
      real*8 dpar(4,*),hpo(14,*),opo(14,*),dparo(4,*),dhpo(14,*)
      real*8 dopo(14,*),deno(mkind,*)
      common /codes/posn(matom,3),kkind(matom),natoms
      common /pairs/tt_list(3,mpair),
     $     dlv_list(3,mpair),dist_list(mpair),
     $     screen_list(mpair),
     $     jkind_list(mpair,2),
     $     jatm_list(mpair,2),
     $     npairs
      logical jsover
      common /lia/ npts,jsover
      common /params/ param(npard,mspin)
      common /refstate/ paros(4,matom),hop(mpkind,mpair),
     $     over(mpkind,mpair),dparos(4,matom),dhop(mpkind,mpair),
     $     dover(mpkind,mpair),den(mkind,matom,mspin)
      real*8 spinos(4,matom),dspnos(4,matom)
      do ip = 1,npairs
         do j = 1,mpkind
            hop(j,ip) = 0d0
            over(j,ip) = 0d0
            dhop(j,ip) = 0d0
            dover(j,ip) = 0d0
         end do
      end do
      call setpar(jspin,param,jsover,den,flipos,.false.,
     $     paros,spinos,hop,over,.false.,ierr)
      if(ierr.ne.0) return
      call setparv(jspin,param,thisvol,jsover,den,flipos,.false.,
     $     dparos,dspnos,dhop,dover,.false.,ierr)
      do ia = 1,natoms
         do k = 1,4
            dpar(k,ia) = paros(k,ia)
            dparo(k,ia) = dparos(k,ia)
         end do
         do k = 1,mkind
            deno(k,ia) = den(k,ia,1)
         end do
      end do
      do ip = 1,npairs
         do j = 1,14
            hpo(j,ip) = hop(j,ip)
            opo(j,ip) = over(j,ip)
            dhpo(j,ip) = dhop(j,ip)
            dopo(j,ip) = dover(j,ip)
         end do
      end do
      return
      end
c
c     Packed H(k), S(k) (ideriv=0) or dH/dV, dS/dV (ideriv=1).
c
      subroutine ref_setham(ak,ideriv,hpk,spk)
     $     bind(C,name='ref_setham')
      implicit real*8 (a-h,o-z)
      include 'P1'
      real*8 ak(3)
      integer ideriv
      complex*16 hpk(*),spk(*)
      common /struc/ nv
      logical jsover
      common /lia/ npts,jsover
      common /refstate/ paros(4,matom),hop(mpkind,mpair),
     $     over(mpkind,mpair),dparos(4,matom),dhop(mpkind,mpair),
     $     dover(mpkind,mpair),den(mkind,matom,mspin)
      complex*16 hmat(mh*(mh+1)/2),smat(mh*(mh+1)/2)
      if(ideriv.eq.0) then
         call setham(jsover,ak,paros,hop,over,hmat,smat,.false.)
      else
         call sethamv(jsover,ak,dparos,dhop,dover,hmat,smat,.false.)
      end if
      do i = 1,nv*(nv+1)/2
         hpk(i) = hmat(i)
         spk(i) = smat(i)
      end do
      return
      end
c
c     ref_setham_nc: the spinor H(k), S(k) of bande.f:897-962 (ideriv 0)
c     or their volume derivatives of bande.f:905-907, 1024-1034 (ideriv
c     1) at one k: spinham on the spins spn (3,natoms) and the on-site
c     spin terms spo (spinos or dspnos of ref_setpar_nc, which must run
c     first), setham / sethamv into the packed nv triangle, and addspins
c     to the packed 2nv triangle, copied out as (2nv)(2nv+1)/2 entries.
      subroutine ref_setham_nc(ak,ideriv,spn,spo,hpk,spk,ierr)
     $     bind(C,name='ref_setham_nc')
      implicit real*8 (a-h,o-z)
      include 'P1'
      real*8 ak(3),spn(3,*),spo(4,*)
      integer ideriv,ierr
      complex*16 hpk(*),spk(*)
      common /struc/ nv
      common /codes/posn(matom,3),kkind(matom),natoms
      common /parinfo/ valence(mkind),kinds,kbas(mkind)
      logical jsover
      common /lia/ npts,jsover
      common /refstate/ paros(4,matom),hop(mpkind,mpair),
     $     over(mpkind,mpair),dparos(4,matom),dhop(mpkind,mpair),
     $     dover(mpkind,mpair),den(mkind,matom,mspin)
      complex*16 hmat(mh*(mh+1)/2),smat(mh*(mh+1)/2)
      real*8 spinat(3,matom),spinos(4,matom)
      real*8 dspth(3,matom),dspphi(3,matom)
      real*8 spndiag(mh/mcol),spndth(mh/mcol)
      complex*16 spnmix(mh/mcol),spnmxth(mh/mcol),spndph(mh/mcol)
      ierr = 0
      if(mcol.lt.2) then
         ierr = 10
         return
      end if
      do i = 1,matom
         do j = 1,3
            spinat(j,i) = 0d0
            dspth(j,i) = 0d0
            dspphi(j,i) = 0d0
         end do
         do j = 1,4
            spinos(j,i) = 0d0
         end do
      end do
      do i = 1,natoms
         do j = 1,3
            spinat(j,i) = spn(j,i)
         end do
         do j = 1,4
            spinos(j,i) = spo(j,i)
         end do
      end do
      call spinham(natoms,kkind,kbas,spinat,spinos,
     $     spndiag,spnmix,.false.,dspth,dspphi,
     $     spndth,spnmxth,spndph)
      if(ideriv.eq.0) then
         call setham(jsover,ak,paros,hop,over,hmat,smat,.false.)
      else
         call sethamv(jsover,ak,dparos,dhop,dover,hmat,smat,.false.)
      end if
      call addspins(nv,spndiag,spnmix,hmat,smat,.false.)
      do i = 1,(2*nv)*(2*nv+1)/2
         hpk(i) = hmat(i)
         spk(i) = smat(i)
      end do
      return
      end
c
c     diag with packed inputs (copied); eigenvectors out as z(nv,nv).
c
      subroutine ref_diag(mode,hpk,spk,eo,zo,ioverr)
     $     bind(C,name='ref_diag')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P2'
      integer mode,ioverr
      complex*16 hpk(*),spk(*),zo(*)
      real*8 eo(*)
      common /struc/ nv
      logical jsover
      common /lia/ npts,jsover
      common /codes/posn(matom,3),kkind(matom),natoms
      common /parinfo/ valence(mkind),kinds,kbas(mkind)
      common /kstuff/ ksk(matom)
      common /refzeig/ zeig(mhe,mhe)
      complex*16 zeig
      complex*16 hmat(mh*(mh+1)/2),smat(mh*(mh+1)/2)
      real*8 en(mh),spev(3,msdir*(matom-1)+1,msdir*(mh-1)+1),ak(3)
      do i = 1,nv*(nv+1)/2
         hmat(i) = hpk(i)
         smat(i) = spk(i)
      end do
      ak(1) = 0d0
      ak(2) = 0d0
      ak(3) = 0d0
      call diag(mode,.false.,hmat,smat,en,.false.,spev,nv,zeig,
     $     jsover,ak,1d3,1d0,ioverr,natoms,kkind,kbas,ksk)
      do i = 1,nv
         eo(i) = en(i)
      end do
      do j = 1,nv
         do i = 1,nv
            zo(i+(j-1)*nv) = zeig(i,j)
         end do
      end do
      return
      end
c
c     perturb on the eigenvectors of the last ref_diag call.
c
      subroutine ref_perturb(dhpk,dspk,eig,eigcut,ntot,deig)
     $     bind(C,name='ref_perturb')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P2'
      complex*16 dhpk(*),dspk(*)
      real*8 eig(*),deig(*),eigcut
      integer ntot
      common /struc/ nv
      common /refzeig/ zeig(mhe,mhe)
      complex*16 zeig
      complex*16 hmat(mh*(mh+1)/2),smat(mh*(mh+1)/2),work(mhp)
      real*8 en(mh),de(mh)
      do i = 1,nv*(nv+1)/2
         hmat(i) = dhpk(i)
         smat(i) = dspk(i)
      end do
      do i = 1,nv
         en(i) = eig(i)
      end do
      call perturb(mhp,nv,hmat,smat,zeig,en,eigcut,ntot,de,work)
      do i = 1,nv
         deig(i) = de(i)
      end do
      return
      end
c
c     fermi on eigen(nv,nk) with weights w(nk).
c
      real*8 function ref_fermi(nk,el,w,eig,tkb,zfill)
     $     bind(C,name='ref_fermi')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P3'
      integer nk
      real*8 w(*),eig(*),el,tkb,zfill
      logical master,multiprc
      common /MPIcomm/ MPISIZE,MPIRANK,master,multiprc
      common /struc/ nv
      common /refeig/ esk(mh,nkdproc),wt(nkd)
      real*8 fermi
      external fermi
      do k = 1,nk
         wt(k) = w(k)
         do i = 1,nv
            esk(i,k) = eig(i+(k-1)*nv)
         end do
      end do
      ref_fermi = fermi(mh,nkd,nkdproc,1,nk,nv,1,el,wt,esk,tkb,zfill)
      return
      end
c
c     The whole bande k-loop on an explicit k list (fractional
c     reciprocal-lattice units) with weights.
c
      subroutine ref_bande(mode,nk,kp,w,thisvol,zfill,tkb,eigcut,
     $     efermi,earg1,earg2,iberr) bind(C,name='ref_bande')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P3'
      integer mode,nk,iberr
      real*8 kp(3,*),w(*),thisvol,zfill,tkb,eigcut
      real*8 efermi,earg1,earg2
      logical master,multiprc
      common /MPIcomm/ MPISIZE,MPIRANK,master,multiprc
      logical jsover
      common /lia/ npts,jsover
      common /cd1/ x(nkd),y(nkd),z(nkd)
      common /refeig/ esk(mh,nkdproc),wt(nkd)
      logical flipos(matom)
      real*8 spinat(3,matom),dspth(3,matom),dspphi(3,matom)
      character*20 label
      iberr = 0
      if(nk.gt.nkd) then
         iberr = 10
         return
      end if
      do i = 1,matom
         flipos(i) = .false.
      end do
      npts = nk
      do k = 1,nk
         x(k) = kp(1,k)
         y(k) = kp(2,k)
         z(k) = kp(3,k)
         wt(k) = w(k)
      end do
      label = 'ref_bande'
      call bande(mode,1,flipos,.false.,spinat,.false.,dspth,dspphi,
     $     .false.,thisvol,zfill,tkb,eigcut,wt,efermi,earg1,earg2,
     $     smag,.false.,.false.,.false.,0,label,iberr)
      return
      end
c
c     The same with jspins = 2 (collinear): param(:,2) from pdn(npar),
c     atom ia flipped where iflip(ia) is non-zero; smag returned.
c
      subroutine ref_bande_spin(mode,nk,kp,w,thisvol,zfill,tkb,eigcut,
     $     pdn,npar,iflip,efermi,earg1,earg2,smag,iberr)
     $     bind(C,name='ref_bande_spin')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P3'
      integer mode,nk,npar,iberr,iflip(*)
      real*8 kp(3,*),w(*),thisvol,zfill,tkb,eigcut,pdn(npar)
      real*8 efermi,earg1,earg2,smag
      logical master,multiprc
      common /MPIcomm/ MPISIZE,MPIRANK,master,multiprc
      logical jsover
      common /lia/ npts,jsover
      common /cd1/ x(nkd),y(nkd),z(nkd)
      common /refeig/ esk(mh,nkdproc),wt(nkd)
      common /codes/posn(matom,3),kkind(matom),natoms
      common /params/ param(npard,mspin)
      logical flipos(matom)
      real*8 spinat(3,matom),dspth(3,matom),dspphi(3,matom)
      character*20 label
      iberr = 0
      if(nk.gt.nkd.or.mspin.lt.2) then
         iberr = 10
         return
      end if
      do i = 1,npard
         param(i,2) = 0d0
      end do
      do i = 1,npar
         param(i,2) = pdn(i)
      end do
      do i = 1,matom
         flipos(i) = .false.
      end do
      do i = 1,natoms
         flipos(i) = iflip(i).ne.0
      end do
      npts = nk
      do k = 1,nk
         x(k) = kp(1,k)
         y(k) = kp(2,k)
         z(k) = kp(3,k)
         wt(k) = w(k)
      end do
      label = 'ref_bande_spin'
      call bande(mode,2,flipos,.false.,spinat,.false.,dspth,dspphi,
     $     .false.,thisvol,zfill,tkb,eigcut,wt,efermi,earg1,earg2,
     $     smag,.false.,.false.,.false.,0,label,iberr)
      return
      end
c
c     The same with lnoncol (non-collinear, jspins = 2): param(:,2) from
c     pdn(npar), spinat(3,natoms) the atoms' unit spin directions.
c
      subroutine ref_bande_nc(mode,nk,kp,w,thisvol,zfill,tkb,eigcut,
     $     pdn,npar,spn,efermi,earg1,earg2,iberr)
     $     bind(C,name='ref_bande_nc')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P3'
      integer mode,nk,npar,iberr
      real*8 kp(3,*),w(*),thisvol,zfill,tkb,eigcut,pdn(npar),spn(3,*)
      real*8 efermi,earg1,earg2
      logical master,multiprc
      common /MPIcomm/ MPISIZE,MPIRANK,master,multiprc
      logical jsover
      common /lia/ npts,jsover
      common /cd1/ x(nkd),y(nkd),z(nkd)
      common /refeig/ esk(mh,nkdproc),wt(nkd)
      common /codes/posn(matom,3),kkind(matom),natoms
      common /params/ param(npard,mspin)
      logical flipos(matom)
      real*8 spinat(3,matom),dspth(3,matom),dspphi(3,matom)
      character*20 label
      iberr = 0
      if(nk.gt.nkd.or.mspin.lt.2.or.mcol.lt.2) then
         iberr = 10
         return
      end if
      do i = 1,npard
         param(i,2) = 0d0
      end do
      do i = 1,npar
         param(i,2) = pdn(i)
      end do
      do i = 1,matom
         flipos(i) = .false.
         do j = 1,3
            spinat(j,i) = 0d0
         end do
      end do
      do i = 1,natoms
         do j = 1,3
            spinat(j,i) = spn(j,i)
         end do
      end do
      npts = nk
      do k = 1,nk
         x(k) = kp(1,k)
         y(k) = kp(2,k)
         z(k) = kp(3,k)
         wt(k) = w(k)
      end do
      label = 'ref_bande_nc'
      call bande(mode,2,flipos,.true.,spinat,.false.,dspth,dspphi,
     $     .false.,thisvol,zfill,tkb,eigcut,wt,efermi,earg1,earg2,
     $     smag,.false.,.false.,.false.,0,label,iberr)
      return
      end
c
c     ref_kptin: the reference's k-point input (kptin, which calls rdspgp
c     and kptgen) on a unit-10 input file fin (the NEWSYM block of a
c     SKIN), with plv from lat.  Unit 15 goes to the file fscr.  Returns
c     the k-points and weights bande receives (/cd1/, /kmesh1/), the
c     lattice (/lattyp/) and the operations (/space/).  File names are
c     passed as byte arrays of the given lengths.
      subroutine ref_kptin(lat,fin,nfin,fscr,nfscr,maxk,nk,qk,wk,
     $     avo,bvo,bijo,volo,lato,nopo,rspo,rspto,tnpo,ierr)
     $     bind(C,name='ref_kptin')
      implicit real*8 (a-h,o-z)
      include 'P1'
      include 'P3'
      integer nfin,nfscr,maxk,nk,lato,nopo,ierr
      integer*1 fin(nfin),fscr(nfscr)
      real*8 lat(3,3),qk(3,maxk),wk(maxk)
      real*8 avo(3,3),bvo(3,3),bijo(3,3),volo
      real*8 rspo(3,3,nopd),rspto(3,3,nopd),tnpo(3,nopd)
      character*255 fn
      common /latt1/plv(3,3)
      logical jsover
      common /lia/ npts,jsover
      common /cd1/ x(nkd),y(nkd),z(nkd)
      common /kmesh1/ qkpt(3,nkd,nwdd),wghtk(nkd,nwdd),nkv(nwdd)
      COMMON /LATTYP/ AVEC(3,3),BVEC(3,3),BIJ(3,3),WSVOL,LATTIC
      COMMON /SPACE / RSPC(3,3,NOPD),RSPCT(3,3,NOPD),TNPSPC(3,NOPD),NOP
      ierr = 0
      do i = 1,3
         do j = 1,3
            plv(j,i) = lat(i,j)
         end do
      end do
      fn = ' '
      do i = 1,min(nfscr,255)
         fn(i:i) = char(fscr(i))
      end do
      open(15,file=fn,status='unknown')
      fn = ' '
      do i = 1,min(nfin,255)
         fn(i:i) = char(fin(i))
      end do
      open(10,file=fn,status='old')
      call kptin
      close(10)
      close(15)
      nk = npts
      if(nk.gt.maxk) then
         ierr = 10
         return
      end if
      do k = 1,nk
         qk(1,k) = x(k)
         qk(2,k) = y(k)
         qk(3,k) = z(k)
         wk(k) = wghtk(k,1)
      end do
      do i = 1,3
         do j = 1,3
            avo(i,j) = avec(i,j)
            bvo(i,j) = bvec(i,j)
            bijo(i,j) = bij(i,j)
         end do
      end do
      volo = wsvol
      lato = lattic
      nopo = nop
      do k = 1,nop
         do j = 1,3
            tnpo(j,k) = tnpspc(j,k)
            do i = 1,3
               rspo(i,j,k) = rspc(i,j,k)
               rspto(i,j,k) = rspct(i,j,k)
            end do
         end do
      end do
      return
      end
c
c     ref_input: the reference's parameter-file reader, input(nparam,
c     jspins), on unit 10 = the file fin (one line: the parameter file's
c     name, as SKIN holds it), unit 15 to /dev/null.  Returns what it
c     leaves in /params/, /parinfo/, /cutcom/, /parcom/ and /lia/.
      subroutine ref_input(fin,nfin,npo,jspo,jsovo,nlto,kindo,kbo,
     $     valo,dnno,scro,paro,ierr) bind(C,name='ref_input')
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer nfin,npo,jspo,jsovo,nlto,kindo,ierr
      integer*1 fin(nfin)
      integer kbo(mkind)
      real*8 valo(mkind),dnno(mint),scro(mint),paro(npard,mspin)
      character*255 fn
      common /parinfo/ valence(mkind),kinds,kbas(mkind)
      common /cutcom/ dnnsq(mint),dnn(mint),
     $     screenl(mint),scrninv(mint)
      common /params/ param(npard,mspin)
      logical realov,harrison
      common /parcom/ nltype,realov,harrison
      logical jsover
      common /lia/ npts,jsover
      ierr = 0
      fn = ' '
      do i = 1,min(nfin,255)
         fn(i:i) = char(fin(i))
      end do
      open(15,file='/dev/null',status='unknown')
      open(10,file=fn,status='old',blank='zero')
      call input(nparam,jspins)
      close(10)
      close(15)
      npo = nparam
      jspo = jspins
      jsovo = 0
      if(jsover) jsovo = 1
      nlto = nltype
      kindo = kinds
      do i = 1,kinds
         kbo(i) = kbas(i)
         valo(i) = valence(i)
      end do
      do i = 1,kinds*(kinds+1)/2
         dnno(i) = dnn(i)
         scro(i) = screenl(i)
      end do
      do j = 1,mspin
         do i = 1,npard
            paro(i,j) = param(i,j)
         end do
      end do
      return
      end
c
c     ref_input1: the reference's structure reader on unit 10 = the file
c     fin (a SKIN structure block from its lattice line on), for kinds
c     with kbin orbitals and valin electrons: setlat and input1, then the
c     conversion of the positions to Cartesian coordinates the main
c     program makes before search2 (static.f:1088-1103).  Returns
c     /latt1/, the volume, the atoms (/codes/, /fazon/), the search
c     widths, zfill, the spin flags and directions, and inpstat.
      subroutine ref_input1(fin,nfin,nkin,kbin,valin,zch,plvo,volo,
     $     nato,ndo,kko,posno,posqo,zfo,iflo,lnco,spo,inpo)
     $     bind(C,name='ref_input1')
      implicit real*8 (a-h,o-z)
      include 'P1'
      integer nfin,nkin,nato,inpo,lnco
      integer*1 fin(nfin)
      integer kbin(nkin),ndo(3),kko(matom),iflo(matom)
      real*8 valin(nkin),zch,plvo(3,3),volo,zfo
      real*8 posno(3,matom),posqo(3,matom),spo(3,matom)
      character*255 fn
      common /codes/posn(matom,3),kkind(matom),natoms
      common /latt1/plv(3,3)
      common /parinfo/ valence(mkind),kinds,kbas(mkind)
      common /fazon/posq(matom,3)
      integer nkdarry(6)
      common/disply/ nkdarry
      logical flipos(matom),lnoncol,ltorque,lspdir
      real*8 spinat(3,matom),dspth(3,matom),dspphi(3,matom),temp(3)
      kinds = nkin
      do i = 1,nkin
         kbas(i) = kbin(i)
         valence(i) = valin(i)
      end do
      fn = ' '
      do i = 1,min(nfin,255)
         fn(i:i) = char(fin(i))
      end do
      open(15,file='/dev/null',status='unknown')
      open(10,file=fn,status='old',blank='zero')
      call input1(zch,zfill,thisvol,inpstat,flipos,
     $     lnoncol,ltorque,lspdir,spinat,dspth,dspphi)
      close(10)
      close(15)
      inpo = inpstat
      if(inpstat.ne.0) return
      do iat = 1,natoms
         do i = 1,3
            tempi = 0d0
            do j = 1,3
               tempi = tempi+plv(j,i)*posn(iat,j)
            end do
            temp(i) = tempi
         end do
         do i = 1,3
            posn(iat,i) = temp(i)
         end do
      end do
      do i = 1,3
         do j = 1,3
            plvo(j,i) = plv(i,j)
         end do
         ndo(i) = nkdarry(i)
      end do
      volo = thisvol
      nato = natoms
      zfo = zfill
      lnco = 0
      if(lnoncol) lnco = 1
      do ia = 1,natoms
         kko(ia) = kkind(ia)
         iflo(ia) = 0
         if(flipos(ia)) iflo(ia) = 1
         do j = 1,3
            posno(j,ia) = posn(ia,j)
            posqo(j,ia) = posq(ia,j)
            spo(j,ia) = spinat(j,ia)
         end do
      end do
      return
      end
