"""Shared workload definitions for the tests: the same detector networks and
event draws, built once for the HIP path (gwfast_amd) and once for the CPU
oracle."""
import numpy as np

from gwfast_amd import catalog, gwfastGlobals as glob

# (name, psd) per workload
NET_C0 = [('Virgo', glob.PSD_ET)]  # one L detector (43.6 deg, 10.5 deg, x-arm 115 deg) with the ET-D ASD
NET_C1 = [('ETS', glob.PSD_ET), ('CE1Id', glob.PSD_CE), ('CE2NM', glob.PSD_CE)]


def oracle_net(sites, earth, fmin=2., fmax=None, compute2arms=True):
    from oracle import detector as OD
    out = []
    for name, psd in sites:
        d = glob.detectors[name]
        out.append(OD.Detector(psd, d['shape'], d['lat'], d['long'], d['xax'], useEarthMotion=earth, fmin=fmin,
                               fmax=fmax, compute2arms=compute2arms))
    return out


def hip_net(wf, sites, earth, fmin=2., fmax=None, compute2arms=True, DutyFactor=None):
    from gwfast_amd.network import DetNet
    from gwfast_amd.signal import GWSignal
    sigs = {}
    for name, psd in sites:
        d = glob.detectors[name]
        sigs[name] = GWSignal(wf, psd_path=psd, detector_shape=d['shape'], det_lat=d['lat'], det_long=d['long'],
                              det_xax=d['xax'], useEarthMotion=earth, fmin=fmin, fmax=fmax,
                              compute2arms=compute2arms, DutyFactor=DutyFactor, verbose=False)
    return DetNet(sigs, verbose=False)


def norm_err(F, Fr):
    """max_ij |F_ij - Fr_ij| / sqrt(Fr_ii Fr_jj) per event, then max."""
    d = np.sqrt(np.abs(np.einsum('iin->in', Fr)))
    return float(np.max(np.abs(F - Fr) / (d[:, None, :] * d[None, :, :])))


def bbh(n, seed):
    return catalog.draw_bbh(n, seed=seed)


def bns(n, seed):
    return catalog.draw_bns(n, seed=seed)
