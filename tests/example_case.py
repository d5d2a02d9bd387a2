"""The reference's default example (example/input.inf) as an oracle Config.

The only numeric known-answer data the reference ships is the max-|V| trace
that example/example.out:17-36 prints for this run (3 significant figures,
mixed-precision build, 4 MPI ranks).  The values below restate the run's
inputs:

  * grid, time and frequency parameters: example/input.inf (nx=ny=nz=384,
    dx=dy=dz=0.5, dt=0.02, nt=1000, xbeg=ybeg=-96, zbeg=-10, na=20, PML,
    fq_min/max/ref = 0.02/2.0/1.0, vcut=1.5, ntdec_r=50);
  * the 1D model: example/lhm.dat (7 layers; air above z=0, because the top
    layer starts at depth 0, src/swpc_3d/m_vmodel_lhm.f90:109-124);
  * the source: example/source.dat, format 'xym0ij' (m_source.f90:514-516):
    an explosion (mxx=myy=mzz=0.8165) of mo=1e15 at (0, 0, 2 km), Kupper
    wavelet with t0=0.1 s and tr=4 s.

report__progress prints, at the start of step it with mod(it,50)=0, the
maxima that kernel__vmax gives after it-1 steps, times UC*M0
(src/swpc_3d/main.f90:119-120, m_report.f90:134-179).
"""
from oracle import oracle as O

NT = 1000
NTDEC_R = 50

LAYERS = [[0.0, 2.3, 5.5, 3.14, 600.0, 300.0],
          [3.0, 2.4, 6.0, 3.55, 600.0, 300.0],
          [18.0, 2.8, 6.7, 3.83, 600.0, 300.0],
          [33.0, 3.2, 7.8, 4.46, 600.0, 300.0],
          [100.0, 3.3, 8.0, 4.57, 600.0, 300.0],
          [225.0, 3.4, 8.4, 4.80, 600.0, 300.0],
          [325.0, 3.5, 8.6, 4.91, 600.0, 300.0]]

SOURCE = dict(x=0.0, y=0.0, z=2.0, t0=0.1, tr=4.0, mo=1e15,
              mxx=0.8165, myy=0.8165, mzz=0.8165, myz=0.0, mxz=0.0, mxy=0.0)


def example_config(nt: int = NT) -> O.Config:
    return O.Config(nx=384, ny=384, nz=384, dx=0.5, dy=0.5, dz=0.5, dt=0.02, nt=nt, na=20, nm=3,
                    xbeg=-96.0, ybeg=-96.0, zbeg=-10.0, fq_min=0.02, fq_max=2.0, fq_ref=1.0,
                    vmodel="lhm", layers=LAYERS, vcut=1.5, sources=[dict(SOURCE)], stftype="kupper")


def report_steps(nt: int = NT):
    """steps it at which report__progress prints; the fields then hold it-1 steps"""
    return [it for it in range(1, nt + 1) if it % NTDEC_R == 0]


def es92(x: float) -> str:
    """Fortran ES9.2 edit of a non-negative value, e.g. 2.94E-05"""
    return f"{x:.2E}"
