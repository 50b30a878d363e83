import numpy as np
deg=5
x=np.cos(np.pi*(np.arange(4000)+0.5)/4000)*0.5
y=2.0**x
w=np.ones_like(x)
for it in range(300):
    # p(f) = 1 + sum_{k>=1} c_k f^k ; minimize rel err: (p - y)/y
    A=np.vander(x,deg+1,increasing=True)[:,1:]/y[:,None]
    b=(y-1)/y
    Aw=A*np.sqrt(w)[:,None]
    c,_,_,_=np.linalg.lstsq(Aw,np.sqrt(w)*b,rcond=None)
    err=np.abs(A@c-b)
    w=w*err; w/=w.sum()
c32=np.concatenate([[1.0],c]).astype(np.float32)
xx=np.linspace(-0.5,0.5,400001).astype(np.float32)
def fma(a,b,c): return (a.astype(np.float64)*b.astype(np.float64)+c.astype(np.float64)).astype(np.float32)
p=np.full_like(xx,c32[5])
for k in range(4,-1,-1):
    p=fma(p,xx,np.float32(c32[k])*np.ones_like(xx))
rel=np.abs(p.astype(np.float64)/2.0**xx.astype(np.float64)-1)
print("max rel err", rel.max(), "p(0)=",p[200000], "max p on [-.5,0]", p[:200001].max())
for v in c32: print("%.9e"%v)
