"""Format-parameter catalogue for the tape formats this backend decodes.

The data file is a subset of the reference's format snapshot
(src/VHSDecode.Core/Formats/format-params.snapshot.json, upstream
oyvindln/vhs-decode commit 43155200): PAL and NTSC cases of VHS, S-VHS,
Betamax and Video8.  Lookup follows FormatCatalog (src/VHSDecode.Core/Formats/
FormatCatalog.cs:46-) keyed by (system, tape_format, tape_speed).
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from functools import lru_cache
from typing import Any, Dict

_DATA = os.path.join(os.path.dirname(__file__), "data", "format_params.json")

COLOR_UNDER_FORMATS = frozenset(
    {"VHS", "VHSHQ", "SVHS", "SVHS_ET", "UMATIC", "UMATIC_HI", "UMATIC_SP", "BETAMAX",
     "BETAMAX_HIFI", "SUPERBETA", "VIDEO8", "HI8", "EIAJ", "VCR", "VCR_LP", "VIDEO2000"})


@dataclass(frozen=True)
class FormatParameters:
    """One (system, tape format, speed) case: the reference's FormatParameterSet
    (src/VHSDecode.Core/Formats/FormatParameterSet.cs)."""

    system: str
    tape_format: str
    tape_speed: str
    sysparams: Dict[str, Any] = field(hash=False, compare=False)
    rfparams: Dict[str, Any] = field(hash=False, compare=False)

    @property
    def parent_system(self) -> str:
        return "PAL" if self.system in ("PAL", "MESECAM") else "NTSC"

    @property
    def is_color_under(self) -> bool:
        return self.tape_format in COLOR_UNDER_FORMATS


@lru_cache(maxsize=1)
def _catalogue():
    with open(_DATA, "r", encoding="utf-8") as handle:
        return json.load(handle)


def available_cases():
    return [(c["system"], c["tape_format"], c["tape_speed"]) for c in _catalogue()["cases"]]


def get_format(system: str = "PAL", tape_format: str = "VHS", tape_speed: str = "sp") -> FormatParameters:
    for case in _catalogue()["cases"]:
        if (case["system"], case["tape_format"], case["tape_speed"]) == (system, tape_format, tape_speed):
            return FormatParameters(system, tape_format, tape_speed, case["sysparams"], case["rfparams"])
    raise KeyError(f"no format parameters for {system}/{tape_format}/{tape_speed}")


def get_laserdisc_format(system: str = "NTSC", lowband: bool = False, deemp_low_mhz: float = 0.0,
                         deemp_high_mhz: float = 0.0, deemp_strength: float = 1.0) -> FormatParameters:
    """FormatCatalog.GetLaserDiscParameters (src/VHSDecode.Core/Formats/FormatCatalog.cs:95-112)
    with the ld command's de-emphasis overrides (Decode/DecodeSession.cs:1350-1372):
    --deemp_low / --deemp_high replace the time constants, --deemp_strength is always set."""
    for case in _catalogue()["ld_cases"]:
        if case["system"] == system and bool(case["lowband"]) == bool(lowband):
            rf = dict(case["rfparams"])
            if deemp_low_mhz > 0 or deemp_high_mhz > 0:
                high_tc, low_tc = rf["video_deemp"][0], rf["video_deemp"][1]
                if deemp_low_mhz > 0:
                    low_tc = 1.0 / (deemp_low_mhz * 1_000_000.0)
                if deemp_high_mhz > 0:
                    high_tc = 1.0 / (deemp_high_mhz * 1_000_000.0)
                rf["video_deemp"] = [high_tc, low_tc]
            rf["video_deemp_strength"] = deemp_strength
            return FormatParameters(system, "LD", "", dict(case["sysparams"]), rf)
    raise KeyError(f"no LaserDisc parameters for {system} (lowband={lowband})")
