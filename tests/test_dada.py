"""The DADA raw-file source: keyword search and header fields against the
reference's own header files (tests/golden/dada), the keyword rules of
ascii_header.c, and a written file read back (CPU)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dada")


def _text(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


def test_reference_headers(dspx):
    # Benchmark/examples/SKA1/Band2/header.dada: 81 complex 10 MHz sub-bands
    h = dspx.dada_parse(_text("ska1_band2_header.dada"))
    assert (h["bandwidth"], h["centre_frequency"], h["nchan"]) == (810.0, 1355.0, 81)
    assert (h["nbit"], h["ndim"], h["npol"], h["tsamp"]) == (8, 2, 2, 0.1)
    assert (h["telescope"], h["source"], h["instrument"]) == ("PKS", "J0437-4715", "CASPSR")
    assert (h["ndat"], h["obs_offset"], h["utc_start"]) == (1024000000000, 0, "2010-04-13-02:05:45")
    assert h["hdr_size"] == 0  # a header-only file: no HDR_SIZE
    # Benchmark/header.dada: real-valued (NDIM 1), inverted band, NCHAN absent
    h = dspx.dada_parse(_text("benchmark_header.dada"))
    assert (h["bandwidth"], h["centre_frequency"], h["nchan"], h["ndim"]) == (-400.0, 1382.0, 1, 1)
    assert (h["tsamp"], h["resolution"]) == (0.00125, 4)
    # Kernel/Formats/ska1/tests/data/mid_8b_data_header.txt: HDR_SIZE 16384,
    # UDP_NCHAN and NBIT_OUT must not be taken for NCHAN and NBIT
    h = dspx.dada_parse(_text("ska1_mid_8b_data_header.txt"))
    assert (h["hdr_size"], h["nchan"], h["nbit"], h["npol"], h["ndim"]) == (16384, 555, 8, 2, 2)
    assert (h["tsamp"], h["centre_frequency"], h["bandwidth"]) == (16.276041667, 59.6736, 379.8368)
    assert (h["resolution"], h["telescope"], h["instrument"]) == (17760, "SKAMid", "MidCBF")


def test_keyword_rules(dspx):
    """A keyword counts at the start of a line followed by a space or tab;
    inside another word, or followed by another character, it does not
    (ascii_header_find)."""
    hdr = "CHAN_BW 5\nBWX 7\nNCHAN_PFB 4096\nBW\t10 # MHz\nNCHAN 81\n"
    assert dspx.ascii_header_get(hdr, "BW") == "10"
    assert dspx.ascii_header_get(hdr, "NCHAN") == "81"
    assert dspx.ascii_header_get(hdr, "FREQ") is None
    # the header's first word is taken whatever follows it
    assert dspx.ascii_header_get("FREQ 1400\n", "FREQ") == "1400"


def test_missing_required_key(dspx):
    hdr = _text("ska1_band2_header.dada").replace(b"UTC_START", b"UTC_STARX")
    with pytest.raises(dspx.DspxError, match="UTC_START"):
        dspx.dada_parse(hdr)
    hdr = _text("ska1_band2_header.dada").replace(b"NDIM 2", b"NDIM 3")
    with pytest.raises(dspx.DspxError):
        dspx.dada_parse(hdr)


def write_dada(path, raw, cfreq, bw, tsamp, nbit, hdr_size=4096, extra=""):
    """A DADA file: ASCII header of hdr_size bytes, then raw TFP samples
    [ndat][nchan][npol][2] (complex, nbit 8 or 16)."""
    ndat, nchan, npol = raw.shape[:3]
    text = ("HDR_VERSION 1.0\nHDR_SIZE %d\nTELESCOPE PKS\nSOURCE J0000+0000\nINSTRUMENT dspx\n"
            "FREQ %r\nBW %r\nNCHAN %d\nNPOL %d\nNBIT %d\nNDIM 2\nTSAMP %r\nUTC_START 2024-01-01-00:00:00\n"
            "OBS_OFFSET 0\n%s" % (hdr_size, cfreq, bw, nchan, npol, nbit, tsamp, extra)).encode()
    assert len(text) < hdr_size
    with open(path, "wb") as f:
        f.write(text + b"\0" * (hdr_size - len(text)))
        f.write(np.ascontiguousarray(raw).tobytes())


@pytest.mark.parametrize("hdr_size", [4096, 16384])
def test_read_written_file(dspx, tmp_path, hdr_size):
    rng = np.random.default_rng(2)
    raw = rng.integers(-128, 128, (1000, 3, 2, 2)).astype(np.int8)
    p = tmp_path / "x.dada"
    write_dada(p, raw, 1400.0, 48.0, 1.0 / 16.0, 8, hdr_size)
    info, data = dspx.read_dada(str(p))
    assert (info["nchan"], info["npol"], info["nbit"], info["hdr_size"]) == (3, 2, 8, hdr_size)
    assert data.shape == (1000, 3, 2, 2)
    assert np.array_equal(np.asarray(data).view(np.int8), raw)
    # NDAT limits the samples taken
    write_dada(p, raw, 1400.0, 48.0, 1.0 / 16.0, 8, hdr_size, extra="NDAT 700\n")
    info, data = dspx.read_dada(str(p))
    assert data.shape[0] == 700
