"""Device-resident measurement pools for timing the accumulator with inputs already in HBM.

torch is used only as the device allocator; the arrays are handed to the library as raw device
pointers (memory kind G4ACC_DEVICE).
"""
from __future__ import annotations

import numpy as np


class DevicePool:
    """All measurements of the pool resident in HBM."""

    def __init__(self, torch, measurements):
        from dca_amd import g4acc
        self.measurements = measurements
        Ms, taus, sites, bands = [], [], [], []
        offs = []
        om = ov = 0
        for meas in measurements:
            o = []
            for s in range(2):
                n = meas.n[s]
                o.append((om, ov, n))
                Ms.append(np.asarray(meas.M[s], np.float64).ravel(order="F"))
                taus.append(meas.tau[s])
                sites.append(meas.site[s])
                bands.append(meas.band[s])
                om += n * n
                ov += n
            offs.append(o)
        dev = torch.device("cuda:0")
        self.M = torch.from_numpy(np.concatenate(Ms)).to(dev)
        self.tau = torch.from_numpy(np.concatenate(taus).astype(np.float64)).to(dev)
        self.site = torch.from_numpy(np.concatenate(sites).astype(np.int32)).to(dev)
        self.band = torch.from_numpy(np.concatenate(bands).astype(np.int32)).to(dev)
        torch.cuda.synchronize()
        pm, pt, ps, pb = (self.M.data_ptr(), self.tau.data_ptr(), self.site.data_ptr(),
                          self.band.data_ptr())
        self.walkers = []
        for meas, o in zip(measurements, offs):
            self.walkers.append(g4acc.device_walker_struct(
                meas.n, [pm + 8 * o[s][0] for s in range(2)], [max(1, o[s][2]) for s in range(2)],
                [pt + 8 * o[s][1] for s in range(2)], [ps + 4 * o[s][1] for s in range(2)],
                [pb + 4 * o[s][1] for s in range(2)], meas.sign))

    def batches(self, size):
        from dca_amd import g4acc
        out = []
        for i in range(0, len(self.walkers) - size + 1, size):
            arr = (g4acc.Walker_c * size)(*self.walkers[i:i + size])
            out.append((arr, size))
        return out
