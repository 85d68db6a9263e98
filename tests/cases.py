"""Workload definitions shared by tests and bench (reference coordinates:
y in [0, yly], walls at y=0 and y=yly)."""
import math

# config 1: 2nd-order channel, Re_bulk=5600 (Re_tau~180), FIMP, RK3, dt=0.005
C1 = dict(nxc=64, nyc=64, nzc=64, xlx=4 * math.pi, yly=2.0, zlz=2 * math.pi,
          xnu=2.0 / 5600.0, dt=0.005, gamma=1.5, ubulk=1.0)

# production step: Re_tau=550, 768x384x576, Lx=2pi, Lz=pi, dt=0.002
C5 = dict(nxc=768, nyc=384, nzc=576, xlx=2 * math.pi, yly=2.0, zlz=math.pi,
          xnu=9.9369e-5, dt=0.002, gamma=1.5, ubulk=1.0)

# small shapes exercising other FFT radices and non-square planes
SMALL = dict(nxc=32, nyc=24, nzc=40, xlx=4 * math.pi, yly=2.0, zlz=2 * math.pi,
             xnu=2.0 / 5600.0, dt=0.005, gamma=1.5, ubulk=1.0)
ODD = dict(nxc=48, nyc=37, nzc=36, xlx=3.0, yly=2.0, zlz=1.7, xnu=1.0 / 1000.0,
           dt=0.004, gamma=1.2, ubulk=1.0)
# full and partial stencil tiles side by side (x: 64 + 32, y: rows of 8 and 4)
TILES = dict(nxc=96, nyc=20, nzc=20, xlx=3.0, yly=2.0, zlz=1.5, xnu=1.0 / 1200.0, dt=0.004,
             gamma=1.5, ubulk=1.0)
# large implicit factor: beta*rdx^2 of order one, exercises the full-circulant path
STIFF = dict(nxc=16, nyc=20, nzc=20, xlx=1.0, yly=2.0, zlz=1.0, xnu=0.05, dt=0.05,
             gamma=1.5, ubulk=1.0)
# free-slip top wall (half-channel style boundary)
FREESLIP = dict(nxc=32, nyc=24, nzc=24, xlx=4.0, yly=1.0, zlz=2.0, xnu=1.0 / 2000.0,
                dt=0.005, gamma=1.5, ubulk=1.0, bc=(0, 0, -1, -2, 0, 0))

# 4th-order channel (src/CFD_4th): Re_tau=180 config 192x128x160, Lx=4pi,
# Lz=4pi/3, nu from Re_tau=0.1538 Re_b^0.887741, dt=0.005
C2_4TH = dict(nxc=192, nyc=128, nzc=160, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3,
              xnu=3.4969e-4, dt=0.005, gamma=1.5, ubulk=1.0, order=4)
SMALL4 = dict(C2_4TH, nxc=32, nyc=24, nzc=24)
# half channel (free-slip top), odd-sized FFT lengths, stronger stretching
HALF4 = dict(nxc=48, nyc=20, nzc=36, xlx=3.0, yly=1.0, zlz=2.0, xnu=1.0 / 1500.0, dt=0.004,
             gamma=2.0, ubulk=1.0, order=4, flow_type=2)

# LES grid of the learned-closure configuration: Re_tau=395 channel filtered
# from 384x192x288 by 4 (x, z) and 2 (y) -> 96x96x72, Lx=2pi, Lz=pi,
# Re_bulk ~ 13750 (nu = 2/Re_b), WALE eddy viscosity
LES96 = dict(nxc=96, nyc=96, nzc=72, xlx=2 * math.pi, yly=2.0, zlz=math.pi,
             xnu=2.0 / 13750.0, dt=0.005, gamma=1.5, ubulk=1.0, les_type=2, filter_type=1)
# the same box at the DNS resolution, with the LES scheme (kernel throughput)
LES384 = dict(LES96, nxc=384, nyc=192, nzc=288, dt=0.002)
