/* CPU model of the pipelined sparse strain sweep, checked against the sequential sweep loop */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#define TOL 0.1
#define EPS 1e-6
typedef struct { double x, y, z, w; } d4;
static int nv, ne; static int (*T)[3]; static int *rig; static double (*X)[3]; static double (*L0)[3];
static double dt = 5e-4;
static uint64_t rs = 88172645463325252ull;
static double urand(void){ rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return (rs >> 11) * (1.0/9007199254740992.0); }
static double dist3(const double* a, const double* b){ double s=0; for(int i=0;i<3;i++){ double d=a[i]-b[i]; s+=d*d;} return sqrt(s); }
static int over(const double* o0, const double* o1, const d4* a0, const d4* a1, double len0, int* deg){
  double c0[3]={o0[0]+dt*a0->x,o0[1]+dt*a0->y,o0[2]+dt*a0->z}, c1[3]={o1[0]+dt*a1->x,o1[1]+dt*a1->y,o1[2]+dt*a1->z};
  double ln=dist3(c0,c1), lo=dist3(o0,o1); *deg=0;
  if (lo>EPS && ln>EPS){ double r=(ln-lo)/lo, s=(ln-len0)/len0; return fabs(s)>TOL || fabs(r)>TOL; }
  *deg=1; return 1; }
static void avg(d4* a, d4* b){ double t[3]={a->x+b->x,a->y+b->y,a->z+b->z}; t[0]*=0.5;t[1]*=0.5;t[2]*=0.5; a->x=b->x=t[0]; a->y=b->y=t[1]; a->z=b->z=t[2]; }
static double (*XO)[3];
/* one element's edges with velocities in a[]; returns written mask, counts */
static int eval_elem(int e, d4* a, int* edges, int* degen){
  int np = T[e][2]>=0?3:2; int w=0; int g;
  const double *o[3]={XO[T[e][0]],XO[T[e][1]],np==3?XO[T[e][2]]:NULL};
  if (over(o[0],o[1],&a[0],&a[1],L0[e][0],&g)){ avg(&a[0],&a[1]); w|=3; (*edges)++; *degen+=g; }
  if (np==3){
    if (over(o[1],o[2],&a[1],&a[2],L0[e][1],&g)){ avg(&a[1],&a[2]); w|=6; (*edges)++; *degen+=g; }
    if (over(o[2],o[0],&a[2],&a[0],L0[e][2],&g)){ avg(&a[2],&a[0]); w|=5; (*edges)++; *degen+=g; }
  }
  return w; }
static int sequential(d4* av, int maxit, int* last_edges, long* degen_sum){
  int niter=0, has=1, edges=0; *degen_sum=0;
  while (has && niter++ < maxit){
    edges=0; int dg=0;
    for (int e=0;e<ne;e++){ if(rig[e]) continue; int np=T[e][2]>=0?3:2; d4 a[3]; for(int k=0;k<np;k++) a[k]=av[T[e][k]];
      int w=eval_elem(e,a,&edges,&dg); for(int k=0;k<np;k++) if(w>>k&1) av[T[e][k]]=a[k]; }
    *degen_sum+=dg; has = edges>0;
  }
  *last_edges=edges; return niter > maxit ? maxit : niter; }

/* ---- schedule ---- */
static int lag_offset = 0, mistake = 0;
static int *la_slot, *la_pos; static int *lev, L, delta, *lp, *slot_elem, *slot_of, (*nslot)[3], (*nlev)[3];
static void schedule(int W0, int CAP){
  lev=calloc(ne,4); int *lv=malloc(4L*nv), *lastlev=malloc(4L*nv), *seen=malloc(4L*nv), *cnt;
  int maxval=0; { int* val=calloc(nv,4); for(int e=0;e<ne;e++){ if(rig[e]) continue; int np=T[e][2]>=0?3:2; for(int k=0;k<np;k++) val[T[e][k]]++; } for(int v=0;v<nv;v++) if(val[v]>maxval) maxval=val[v]; free(val);}
  int W = W0>0?W0:(maxval>1?maxval-1:1); int ok=0;
  for(; W<=64 && !ok; W*=2){
    memset(lev,0,4L*ne); for(int v=0;v<nv;v++) lastlev[v]=-1;
    for(int pass=0;pass<16;pass++){
      int changed=0; for(int v=0;v<nv;v++){lv[v]=-1; seen[v]=0;}
      for(int e=0;e<ne;e++){ if(rig[e]) continue; int np=T[e][2]>=0?3:2; int l=lev[e];
        for(int k=0;k<np;k++){int v=T[e][k]; if(lv[v]+1>l) l=lv[v]+1; if(!seen[v] && lastlev[v]-W>l) l=lastlev[v]-W;}
        if(l!=lev[e]){changed=1; lev[e]=l;}
        for(int k=0;k<np;k++){int v=T[e][k]; lv[v]=l; seen[v]=1;}}
      for(int v=0;v<nv;v++) lastlev[v]=lv[v];
      if(!changed){ ok=1; break; }
    }
    if (ok) printf("relaxation W=%d converged\n", W);
  }
  if(!ok){ /* ASAP */ for(int v=0;v<nv;v++) lv[v]=-1; for(int e=0;e<ne;e++){ if(rig[e]) continue; int np=T[e][2]>=0?3:2; int l=0; for(int k=0;k<np;k++) if(lv[T[e][k]]+1>l) l=lv[T[e][k]]+1; lev[e]=l; for(int k=0;k<np;k++) lv[T[e][k]]=l; } printf("ASAP fallback\n"); }
  L=0; for(int e=0;e<ne;e++) if(!rig[e] && lev[e]+1>L) L=lev[e]+1;
  /* components: union-find over vertices */
  int* uf=malloc(4L*nv); for(int v=0;v<nv;v++) uf[v]=v;
  #define FIND(x) ({int _r=(x); while(uf[_r]!=_r){uf[_r]=uf[uf[_r]]; _r=uf[_r];} _r;})
  for(int e=0;e<ne;e++){ if(rig[e]) continue; int np=T[e][2]>=0?3:2; int r0=FIND(T[e][0]); for(int k=1;k<np;k++){int r=FIND(T[e][k]); if(r!=r0){ if(r<r0){uf[r0]=r; r0=r;} else uf[r]=r0; }}}
  int *cmin=malloc(4L*nv), *cmax=malloc(4L*nv), *csz=calloc(nv,4);
  for(int v=0;v<nv;v++){cmin[v]=1<<30; cmax[v]=-1;}
  for(int e=0;e<ne;e++){ if(rig[e]) continue; int r=FIND(T[e][0]); if(lev[e]<cmin[r]) cmin[r]=lev[e]; if(lev[e]>cmax[r]) cmax[r]=lev[e]; csz[r]++; }
  cnt=calloc(L,4);
  /* fixed (long) components first */
  int nsmall=0; int* small=malloc(4L*nv);
  for(int v=0;v<nv;v++) if(csz[v]>0){ if(2*(cmax[v]-cmin[v]+1) > L) {} else small[nsmall++]=v; }
  for(int e=0;e<ne;e++){ if(rig[e]) continue; int r=FIND(T[e][0]); if(2*(cmax[r]-cmin[r]+1) > L) cnt[lev[e]]++; }
  /* per small component: level histogram relative to its min */
  int* shift=calloc(nv,4);
  int** hist=malloc(sizeof(int*)*nv); memset(hist,0,sizeof(int*)*nv);
  for(int i=0;i<nsmall;i++){ int r=small[i]; hist[r]=calloc(cmax[r]-cmin[r]+1,4);} 
  for(int e=0;e<ne;e++){ if(rig[e]) continue; int r=FIND(T[e][0]); if(hist[r]) hist[r][lev[e]-cmin[r]]++; }
  for(int i=0;i<nsmall;i++){ int r=small[i]; int len=cmax[r]-cmin[r]+1; int range=L-len; int step=range/128>0?range/128:1; int best=-1, bestc=1<<30;
    for(int o=0;o<=range;o+=step){ int m=0; for(int j=0;j<len;j++){ int c=cnt[o+j]+hist[r][j]; if(c>m) m=c; } if(m<bestc){bestc=m; best=o;} }
    shift[r]=best-cmin[r]; for(int j=0;j<len;j++) cnt[best+j]+=hist[r][j]; }
  for(int e=0;e<ne;e++){ if(rig[e]) continue; int r=FIND(T[e][0]); if(hist[r]) lev[e]+=shift[r]; }
  /* split levels above CAP */
  memset(cnt,0,4L*L); for(int e=0;e<ne;e++) if(!rig[e]) cnt[lev[e]]++;
  int* newbase=malloc(4L*(L+1)); int nl=0; for(int l=0;l<L;l++){ newbase[l]=nl; nl += cnt[l]>CAP ? (cnt[l]+CAP-1)/CAP : 1; } 
  int* fill=calloc(L,4); for(int e=0;e<ne;e++){ if(rig[e]) continue; int l=lev[e]; lev[e]=newbase[l]+fill[l]/CAP; fill[l]++; }
  L=nl;
  /* slots */
  lp=calloc(L+1,4); for(int e=0;e<ne;e++) if(!rig[e]) lp[lev[e]+1]++; for(int l=0;l<L;l++) lp[l+1]+=lp[l];
  int nsl=lp[L]; slot_elem=malloc(4L*nsl); slot_of=malloc(4L*ne); int* f2=malloc(4L*L); memcpy(f2,lp,4L*L);
  for(int e=0;e<ne;e++){ slot_of[e]=-1; if(rig[e]) continue; int s=f2[lev[e]]++; slot_elem[s]=e; slot_of[e]=s; }
  /* spans, next visits (cyclic) */
  int *fi=malloc(4L*nv), *fpos=malloc(4L*nv), *la=malloc(4L*nv), *lpos=malloc(4L*nv); for(int v=0;v<nv;v++){fi[v]=-1; la[v]=-1;}
  nslot=malloc(12L*nsl); nlev=malloc(12L*nsl);
  for(int s=0;s<nsl;s++){ nslot[s][0]=nslot[s][1]=nslot[s][2]=-1; }
  for(int e=0;e<ne;e++){ if(rig[e]) continue; int np=T[e][2]>=0?3:2; int s=slot_of[e]; for(int k=0;k<np;k++){ int v=T[e][k];
     if(fi[v]<0){fi[v]=s; fpos[v]=k;} else { nslot[la[v]][lpos[v]]=4*s+k; nlev[la[v]][lpos[v]]=lev[e]; }
     la[v]=s; lpos[v]=k; } }
  la_slot=malloc(4L*nv); la_pos=malloc(4L*nv); for(int v=0;v<nv;v++){ la_slot[v]=la[v]; la_pos[v]=lpos[v]; }
  int mx=0; for(int v=0;v<nv;v++) if(fi[v]>=0){ nslot[la[v]][lpos[v]]=4*fi[v]+fpos[v]; nlev[la[v]][lpos[v]]=lev[slot_elem[fi[v]]]; int sp=lev[slot_elem[la[v]]]-lev[slot_elem[fi[v]]]; if(sp>mx) mx=sp; }
  delta=mx+1; int mxc=0; for(int l=0;l<L;l++) if(lp[l+1]-lp[l]>mxc) mxc=lp[l+1]-lp[l];
  printf("schedule: L=%d slots=%d maxlevel=%d delta=%d\n", L, nsl, mxc, delta);
}

/* ---- pipelined model ---- */
#define R 64
#define CAPL 1024
static long far_marks = 0;
static int pipelined(d4* av, int maxit, int* last_edges, long* degen_sum, long* iters, long* evals){
  int nsl=lp[L];
  d4 (*cp)[3]=malloc(sizeof(d4)*3*nsl); /* local copies, w = origin */
  for(int s=0;s<nsl;s++){ int e=slot_elem[s]; int np=T[e][2]>=0?3:2; for(int k=0;k<np;k++){ cp[s][k]=av[T[e][k]]; cp[s][k].w=-1; } }
  *degen_sum=0; *iters=0; *evals=0; int total_sweeps=0; *last_edges=0;
  int k0=0;
  while (k0 < maxit){
    const int dl = delta + lag_offset > 0 ? delta + lag_offset : 1;
    int K = 64/delta; if (K>10) K=10; if (K<1) K=1; if (K>maxit-k0) K=maxit-k0;
    /* pre-check */
    uint8_t* pre=calloc(nsl,1); int anypre=0;
    for(int s=0;s<nsl;s++){ int e=slot_elem[s]; int np=T[e][2]>=0?3:2; d4 a[3]; for(int k=0;k<np;k++) a[k]=av[T[e][k]]; int ed=0,dg=0; eval_elem(e,a,&ed,&dg); /* note: pre-check in kernel tests each edge on the pre-sweep velocities separately */ pre[s]=0;
      { int g; const double* o[3]={XO[T[e][0]],XO[T[e][1]],np==3?XO[T[e][2]]:NULL}; d4 b[3]; for(int k=0;k<np;k++) b[k]=av[T[e][k]]; int any=over(o[0],o[1],&b[0],&b[1],L0[e][0],&g); if(np==3){ any|=over(o[1],o[2],&b[1],&b[2],L0[e][1],&g); any|=over(o[2],o[0],&b[2],&b[0],L0[e][2],&g);} pre[s]=any; anypre|=any; } }
    if (!anypre){ total_sweeps++; *last_edges=0; break; }
    /* marks R or more levels ahead (one sweep in flight): per-slot bytes,
       taken by the first sweep's scan of that level */
    uint8_t* far=calloc(nsl,1);
    /* levels the first sweep must visit: a pre-check hit or a far mark */
    uint8_t* preb=calloc(L,1);
    for(int s=0;s<nsl;s++) if(pre[s]) preb[lev[slot_elem[s]]]=1;
    uint8_t ring[R][CAPL]; memset(ring,0,sizeof ring); int owner[R]; for(int i=0;i<R;i++) owner[i]=-1;
    long edges[16]={0};
    int tend = L + (K-1)*dl;
    /* as the kernel: start at the first such level; after an iteration, when
       no mark is pending (every mark is consumed within dl iterations of the
       iteration that made it), jump to the next such level */
    int last_mark = -(1<<29);
    int t = 0; while (t < L && !preb[t]) t++;
    for (; t<tend; ){
      /* scan */
      int nw=0; static int wk_s[16*CAPL], wk_k[16*CAPL], wk_f[16*CAPL];
      for (int k=0;k<K;k++){ int l=t-k*dl; if(l<0||l>=L) continue; int n=lp[l+1]-lp[l];
        for(int i=0;i<n;i++){ int f=ring[l&(R-1)][i]; if(k==0){ if(pre[lp[l]+i]) f|=8; f|=far[lp[l]+i]; far[lp[l]+i]=0; } if(f){ wk_s[nw]=lp[l]+i; wk_k[nw]=k; wk_f[nw]=f; nw++; } ring[l&(R-1)][i]=0; }
        if (owner[l&(R-1)]!=-1 && owner[l&(R-1)]!=l){ printf("ring alias at t=%d l=%d owner=%d\n",t,l,owner[l&(R-1)]); exit(1);} owner[l&(R-1)]=-1; }
      if (nw) (*iters)++;
      *evals+=nw;
      /* eval (entries are independent: process in any order; use reverse to catch ordering bugs) */
      for (int q=nw-1;q>=0;q--){
        int s=wk_s[q], k=wk_k[q], f=wk_f[q]; int e=slot_elem[s]; int np=T[e][2]>=0?3:2; int l=lev[e];
        d4 a[3]; for(int p=0;p<np;p++) a[p]=cp[s][p];
        int ed=0, dg=0; int w=eval_elem(e,a,&ed,&dg);
        edges[k]+=ed; *degen_sum+=dg;
        /* A changed vertex's value goes to its next visit, with the id of
           the slot that wrote it.  The next visit is marked (evaluated)
           unless it is that writer, which is evaluated anyway because it
           wrote; the writer's own copy is refreshed by that delivery, so
           it only stores its own copy when no delivery will come: a vertex
           with one visit, or the vertex's last visit in the last sweep. */
        for(int p=0;p<np;p++){
          double origin;
          const int wrote = w>>p&1;
          if (wrote) origin=s; else if ((f>>p&1) && mistake!=1){ origin=cp[s][p].w; } else continue;
          int ns=nslot[s][p]; int ts=ns>>2, tpos=ns&3, tl=nlev[s][p];
          int wrap = tl < l;
          if (ts==s || (wrap && k==K-1)) { if (wrote) { d4 c=a[p]; c.w=s; cp[s][p]=c; } continue; }
          d4 c=a[p]; c.w=origin; cp[ts][tpos]=c;
          if (ts==(int)origin) continue;
          if (!wrap && tl - l >= R){ if (K!=1){ printf("far mark with %d sweeps in flight\n",K); exit(1);} far[ts] |= 1<<tpos; preb[tl]=1; far_marks++; continue; }
          int ri=tl&(R-1); if(owner[ri]!=-1 && owner[ri]!=tl){ printf("ring alias mark t=%d tl=%d owner=%d\n",t,tl,owner[ri]); exit(1);} owner[ri]=tl;
          ring[ri][ts-lp[tl]] |= 1<<tpos;
          last_mark = t;
        }
        if (ed && k<K-1){ int ri=l&(R-1); if(owner[ri]!=-1 && owner[ri]!=l){ printf("ring alias fired\n"); exit(1);} owner[ri]=l; ring[ri][s-lp[l]] |= 8; last_mark = t; }
      }
      int tn = t + 1;
      /* planted mistakes: 2 skips one iteration early, 3 skips with marks pending */
      if (mistake == 3 || t >= last_mark + dl - (mistake == 2 ? 1 : 0)) {
        tn = t + 1; while (tn < L && !preb[tn]) tn++;
        if (tn >= L) break;
      }
      t = tn;
    }
    /* writeback: each vertex's value from its last visit's copy */
    for(int v=0;v<nv;v++) if(la_slot[v]>=0){ d4 c=cp[la_slot[v]][la_pos[v]]; av[v].x=c.x; av[v].y=c.y; av[v].z=c.z; }
    int n=0; while(n<K && edges[n]>0) n++; int niter = n<K ? n+1 : K;
    total_sweeps += niter; *last_edges = edges[niter-1];
    free(pre); free(far); free(preb);
    if (niter < K || edges[K-1]==0) break;
    k0 += K;
    for(int s=0;s<nsl;s++){ int e=slot_elem[s]; int np=T[e][2]>=0?3:2; for(int k=0;k<np;k++){ cp[s][k]=av[T[e][k]]; cp[s][k].w=-1; } }
  }
  free(cp);
  return total_sweeps; }

int main(int argc, char** argv){
  FILE* fp=fopen(argv[1],"rb"); int hdr[2]; if(fread(hdr,4,2,fp)!=2) return 1; nv=hdr[0]; ne=hdr[1];
  T=malloc(12L*ne); rig=malloc(4L*ne); X=malloc(24L*nv); L0=malloc(24L*ne);
  if(fread(T,12,ne,fp)!=(size_t)ne || fread(rig,4,ne,fp)!=(size_t)ne || fread(X,24,nv,fp)!=(size_t)nv || fread(L0,24,ne,fp)!=(size_t)ne) return 1;
  fclose(fp);
  double xs=atof(argv[2]), vs=atof(argv[3]); int maxit=atoi(argv[4]); int cap=argc>5?atoi(argv[5]):1024; int seed=argc>6?atoi(argv[6]):1; lag_offset=argc>7?atoi(argv[7]):0; mistake=argc>8?atoi(argv[8]):0;
  rs += seed*7919;
  schedule(0, cap);
  XO=malloc(24L*nv); d4* av0=malloc(sizeof(d4)*nv);
  /* position noise and velocities; a disc of strong velocity noise */
  for(int v=0;v<nv;v++){ for(int i=0;i<3;i++) XO[v][i]=X[v][i]+xs*(urand()-0.5); double r2=X[v][0]*X[v][0]+X[v][1]*X[v][1]; double sc= r2<0.1? vs : vs*0.01; av0[v].x=sc*(urand()-0.5); av0[v].y=sc*(urand()-0.5); av0[v].z=-1.0+sc*(urand()-0.5); av0[v].w=0; }
  d4* a1=malloc(sizeof(d4)*nv); d4* a2=malloc(sizeof(d4)*nv); memcpy(a1,av0,sizeof(d4)*nv); memcpy(a2,av0,sizeof(d4)*nv);
  int e1,e2; long g1,g2,it,ev; int n1=sequential(a1,maxit,&e1,&g1); int n2=pipelined(a2,maxit,&e2,&g2,&it,&ev);
  long diff=0; for(int v=0;v<nv;v++) if(a1[v].x!=a2[v].x||a1[v].y!=a2[v].y||a1[v].z!=a2[v].z) diff++;
  printf("seq: sweeps=%d edges=%d degen=%ld | pipe: sweeps=%d edges=%d degen=%ld iters=%ld evals=%ld far=%ld | vertices differing=%ld\n", n1,e1,g1,n2,e2,g2,it,ev,far_marks,diff);
  return diff!=0 || n1!=n2 || e1!=e2 || g1!=g2; }
