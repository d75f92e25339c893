"""Regenerate ``spectrum_tables.inc`` from the reference's published data tables.

The tables are data, not code: the CIE 1931 2-degree colour-matching functions
sampled at 1 nm over 360-830 nm (471 samples) and Smits' RGB-to-spectrum basis
spectra (32 samples over 380-720 nm), as held in
``WoodenPBREngine/SpectrumData.h:5-958``.  The reference resamples them into
64 bins over 400-700 nm at start-up (``SampledSpectrum::init``,
``CSpectrumSampled.h:236-281``); that resampling is re-implemented in
``spectrum.cpp`` and is not done here.

Usage (only where the reference checkout exists):
    python gen_spectrum_tables.py /root/reference/WoodenPBREngine/SpectrumData.h
"""
import re
import sys
from pathlib import Path

NAMES = [
    "CIE_X", "CIE_Y", "CIE_Z", "CIE_lambda", "RGB2SpectLambda",
    "RGBRefl2SpectWhite", "RGBRefl2SpectCyan", "RGBRefl2SpectMagenta",
    "RGBRefl2SpectYellow", "RGBRefl2SpectRed", "RGBRefl2SpectGreen",
    "RGBRefl2SpectBlue",
    "RGBIllum2SpectWhite", "RGBIllum2SpectCyan", "RGBIllum2SpectMagenta",
    "RGBIllum2SpectYellow", "RGBIllum2SpectRed", "RGBIllum2SpectGreen",
    "RGBIllum2SpectBlue",
]
SIZES = {"CIE_X": 471, "CIE_Y": 471, "CIE_Z": 471, "CIE_lambda": 471}


def parse(text):
    out = {}
    for name in NAMES:
        m = re.search(r"const\s+float\s+%s\s*\[[^\]]*\]\s*=\s*\{(.*?)\};" % name,
                      text, re.S)
        if not m:
            raise SystemExit("table %s not found" % name)
        body = re.sub(r"//[^\n]*", "", m.group(1))
        vals = [v.strip().rstrip("fF") for v in body.split(",") if v.strip()]
        want = SIZES.get(name, 32)
        if len(vals) != want:
            raise SystemExit("table %s has %d values, expected %d" % (name, len(vals), want))
        out[name] = [float(v) for v in vals]
    m = re.search(r"CIE_Y_integral\s*=\s*([0-9.eE+-]+)", text)
    out["CIE_Y_integral"] = float(m.group(1))
    return out


def lit(v):
    t = "%.9g" % v
    if not any(ch in t for ch in ".eEn"):
        t += ".0"
    return t + "f"


def emit(tables):
    lines = ["// Generated by gen_spectrum_tables.py from the reference's SpectrumData.h",
             "// (CIE 1931 2-degree observer, 1 nm; Smits RGB->spectrum bases). Data only.",
             "#pragma once",
             "#define WPBR_N_CIE 471",
             "#define WPBR_N_RGB2SPECT 32",
             "static const float WPBR_CIE_Y_INTEGRAL = %s;" % lit(tables["CIE_Y_integral"])]
    for name in NAMES:
        vals = tables[name]
        lines.append("static const float WPBR_%s[%d] = {" % (name, len(vals)))
        for i in range(0, len(vals), 6):
            lines.append("  " + ", ".join(lit(v) for v in vals[i:i + 6]) + ",")
        lines.append("};")
    return "\n".join(lines) + "\n"


def main():
    src = Path(sys.argv[1] if len(sys.argv) > 1 else
               "/root/reference/WoodenPBREngine/SpectrumData.h")
    tables = parse(src.read_text(encoding="utf-8", errors="replace"))
    dst = Path(__file__).with_name("spectrum_tables.inc")
    dst.write_text(emit(tables))
    print("wrote", dst)


if __name__ == "__main__":
    main()
