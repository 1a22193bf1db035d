"""Case configuration: the ofx_config struct of include/ofx.h and the reference cases.

Values of the three reference regression cases are those of their script files
(test/<case>/script/{cfd,inflow,turb,vismodel}.txt) with the defaults the
reference applies to keys the scripts leave out.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass


class OfxConfig(ctypes.Structure):
    """Mirror of `ofx_config` in include/ofx.h (field order and types must match)."""

    _fields_ = [
        ("dim", ctypes.c_int32),
        ("ischeme", ctypes.c_int32),
        ("ilim", ctypes.c_int32),
        ("ieigenfix", ctypes.c_int32),
        ("ivischeme", ctypes.c_int32),
        ("vismodel", ctypes.c_int32),
        ("timestepModel", ctypes.c_int32),
        ("visSRModel", ctypes.c_int32),
        ("visTimeStepModel", ctypes.c_int32),
        ("ncfl", ctypes.c_int32),
        ("isowallbc", ctypes.c_int32),
        ("iturb_visflux", ctypes.c_int32),
        ("iprod_sa", ctypes.c_int32),
        ("turb_ilim", ctypes.c_int32),
        ("tns_ilim", ctypes.c_int32),
        ("flux_fp32", ctypes.c_int32),
        ("time_integral", ctypes.c_int32),
        ("rk_stage", ctypes.c_int32),
        ("vencat_coef", ctypes.c_double),
        ("centropy1", ctypes.c_double),
        ("centropy2", ctypes.c_double),
        ("cflst", ctypes.c_double),
        ("cfled", ctypes.c_double),
        ("max_time_ratio", ctypes.c_double),
        ("mach_ref", ctypes.c_double),
        ("aoa_degree", ctypes.c_double),
        ("sideslip_degree", ctypes.c_double),
        ("reynolds", ctypes.c_double),
        ("gama_ref", ctypes.c_double),
        ("prl", ctypes.c_double),
        ("prt", ctypes.c_double),
        ("tref_dim", ctypes.c_double),
        ("twall_dim", ctypes.c_double),
        ("ref_sa", ctypes.c_double),
        ("max_vis_ratio", ctypes.c_double),
        ("rprod", ctypes.c_double),
        ("turb_cfl_ratio", ctypes.c_double),
        ("aref", ctypes.c_double),
    ]


@dataclass
class CaseConfig:
    """Python view of the case keys; `to_c()` packs the C struct."""

    dim: int = 2
    ischeme: int = 1
    ilim: int = 0
    ieigenfix: int = 1
    ivischeme: int = 0
    vismodel: int = 1
    timestepModel: int = 0
    visSRModel: int = 2
    visTimeStepModel: int = 1
    ncfl: int = 100
    isowallbc: int = 0
    iturb_visflux: int = 2
    iprod_sa: int = 2
    turb_ilim: int = -1
    tns_ilim: int = -1
    flux_fp32: int = 0
    time_integral: int = 2
    rk_stage: int = 1
    vencat_coef: float = 0.0001
    centropy1: float = 0.0001
    centropy2: float = 0.0001
    cflst: float = 0.1
    cfled: float = 10.0
    max_time_ratio: float = 1.0e10
    mach_ref: float = 0.1
    aoa_degree: float = 0.0
    sideslip_degree: float = 0.0
    reynolds: float = 2.0e5
    gama_ref: float = 1.4
    prl: float = 0.72
    prt: float = 0.90
    tref_dim: float = 300.0
    twall_dim: float = 300.0
    ref_sa: float = 0.1
    max_vis_ratio: float = 1.0e6
    rprod: float = 1.0
    turb_cfl_ratio: float = 1.0
    aref: float = 1.0

    def to_c(self) -> OfxConfig:
        c = OfxConfig()
        for name, _ in OfxConfig._fields_:
            setattr(c, name, getattr(self, name))
        return c

    def replace(self, **kw) -> "CaseConfig":
        d = dict(self.__dict__)
        d.update(kw)
        return CaseConfig(**d)


# test/plateuns2dslau2/script/{cfd,inflow,vismodel}.txt: laminar flat plate, SLAU2
PLATE_SLAU2 = CaseConfig(ischeme=12, ilim=0, ivischeme=0, vismodel=1, mach_ref=0.1, reynolds=2.0e5,
                         centropy1=0.0001, centropy2=0.0001, vencat_coef=0.0001)
# test/turbplateuns2droe_sa/script/*: turbulent flat plate, Roe + SA
TURBPLATE_ROE_SA = CaseConfig(ischeme=1, ilim=0, ivischeme=0, vismodel=3, mach_ref=0.2, reynolds=1.03e7,
                              centropy1=0.0001, centropy2=0.0001, vencat_coef=0.0001)
# test/rae2822_roe_sa/script/*: RAE2822, Roe + SA + Venkatakrishnan
RAE2822_ROE_SA = CaseConfig(ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=0.729, aoa_degree=2.31,
                            reynolds=6.5e6, centropy1=0.001, centropy2=0.001, vencat_coef=0.001)

REFERENCE_CASES = {
    "plateuns2dslau2": PLATE_SLAU2,
    "turbplateuns2droe_sa": TURBPLATE_ROE_SA,
    "rae2822_roe_sa": RAE2822_ROE_SA,
}
