"""Time averages of the plane statistics, as TAVER keeps them
(compact3d.f:10755-11072).

TAVER(NAVG) is called once per averaging interval after ENERGY (MAIN,
compact3d.f:113-117).  On the first call (NAVG = 1) it copies each
instantaneous quantity X into its average XO (:10809-10937); afterwards
XO = (XO*REAL(NAVG-1) + X)/REAL(NAVG) (:10939-11069).  TimeAverage applies
the same rule to the quantities this library forms: ENERGY's plane
statistics (Channel.stats(): VM, PM, WM, VRMS, PRMS, PV, WRMS, U2V, VW2),
WSM(1:3) -> WSMO, TAOW -> TAOWO, and the wall stresses TAUW/TAUB -> TAUWO/
TAUBO.  The spectra, correlations, quadrant and budget terms that ENERGY
also forms are not computed by the library, so they are not averaged here.
"""
from __future__ import annotations

import numpy as np


class TimeAverage:
    def __init__(self):
        self.navg = 0
        self.avg = {}

    def _acc(self, name, x):
        x = np.array(x, dtype=np.float64, copy=True)
        if self.navg == 1 or name not in self.avg:
            self.avg[name] = x
        else:
            n = float(self.navg)
            self.avg[name] = (self.avg[name] * (n - 1.0) + x) / n

    def add(self, stats=None, wsm=None, taow=None, tauw=None, taub=None):
        """one TAVER call: NAVG += 1, then every quantity given is folded
        into its average.  stats: Channel.stats() (20, N2-1); wsm: WSM(1:3);
        taow: TAOW; tauw/taub: the wall stress arrays."""
        self.navg += 1
        for name, x in (("stats", stats), ("wsmo", wsm), ("taowo", taow), ("tauwo", tauw), ("taubo", taub)):
            if x is not None:
                self._acc(name, x)
        return self.navg

    def __getitem__(self, name):
        return self.avg[name]

    def profile(self, name):
        """the averaged profile (J = 1..N2-1) of one statistic by its name
        in Channel.STAT_NAMES, e.g. 'VM1' (the mean streamwise velocity)"""
        from .c3d import Channel
        return self.avg["stats"][Channel.STAT_NAMES.index(name)]
