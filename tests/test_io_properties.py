"""Property tests of the host-side formats (CPU, hypothesis): generated
config.yaml texts in block and flow style, with and without quotes and
comments, read by the C loader (mcr_load_config) against the values written
(PyYAML checks the text is YAML with those keys) and Python's own number
parsing for the scalars; results-CSV rows with arbitrary comments and index
names read back by Python's csv module; arbitrary text into the config
loader and arbitrary bytes into the binary readers, which must answer with a
result or an error code (tests/sanitize.sh runs these under ASan/UBSan).

What pins the loader's semantics is the reference's load_config
(src/utilities/load_config.c:11-25 scalar conversions by atoi/atol/atof,
:207-343 keys, sequences and mappings) and its CSV writer
(src/utilities/csv_writer.c:282-315: fields holding ',', '"', CR or LF quoted,
'"' doubled)."""
from __future__ import annotations

import csv
import io
import os

import pytest

hyp = pytest.importorskip("hypothesis")
yaml = pytest.importorskip("yaml")
from hypothesis import HealthCheck, given, settings  # noqa: E402
from hypothesis import strategies as st  # noqa: E402

# MCR_HYP_EXAMPLES raises the number of generated cases for a longer run
SETTINGS = settings(max_examples=int(os.environ.get("MCR_HYP_EXAMPLES", "150")), deadline=None,
                    suppress_health_check=[HealthCheck.function_scoped_fixture])

ident = st.from_regex(r"[a-z][a-z0-9_]{0,10}", fullmatch=True)
index_name = st.from_regex(r"[A-Z][A-Za-z0-9]{0,8}", fullmatch=True)
date = st.dates().map(lambda d: d.strftime("%Y-%m-%d"))
# text a double-quoted YAML scalar and the loader both keep verbatim
comment_text = st.text(alphabet=st.characters(min_codepoint=32, max_codepoint=126,
                                              blacklist_characters='"\\'),
                       max_size=40)


def q(s, how):
    return {"none": s, "single": f"'{s}'", "double": f'"{s}"'}[how]


quoting = st.sampled_from(["none", "single", "double"])


@st.composite
def configs(draw):
    c = {
        "models": draw(st.lists(ident, min_size=0, max_size=5)),
        "indices": draw(st.dictionaries(index_name, ident.map(lambda s: s + ".csv"),
                                        max_size=5)),
        "start": draw(st.one_of(st.none(), date)),
        "end": draw(st.one_of(st.none(), date)),
        "x": draw(st.one_of(st.none(), st.floats(0.0, 1.0, allow_nan=False))),
        "M": draw(st.one_of(st.none(), st.integers(1, 2**31 - 1))),
        "k": draw(st.one_of(st.none(), st.integers(0, 1000))),
        "N": draw(st.one_of(st.none(), st.integers(0, 10**6))),
        "random_seed": draw(st.one_of(st.none(), st.integers(0, 2**32 - 1))),
        "train_ratio": draw(st.one_of(st.none(), st.floats(0.0, 1.0, allow_nan=False))),
        "comment": draw(st.one_of(st.none(), comment_text)),
    }
    style = {"models": draw(st.sampled_from(["block", "flow"])),
             "indices": draw(st.sampled_from(["block", "flow"]))}
    quotes = draw(st.lists(quoting, min_size=16, max_size=16))
    trailing = draw(st.booleans())
    order = draw(st.permutations(list(c)))
    return c, style, quotes, trailing, order


def render(c, style, quotes, trailing, order):
    qi = iter(quotes * 4)
    lines = ["# generated configuration"]
    tail = "  # note" if trailing else ""
    for key in order:
        v = c[key]
        if v is None:
            continue
        if key == "models":
            if style["models"] == "flow":
                lines.append("models: [" + ", ".join(q(m, next(qi)) for m in v) + "]" + tail)
            else:
                lines.append("models:" + tail)
                lines += [f"  - {q(m, next(qi))}" for m in v]
        elif key == "indices":
            if style["indices"] == "flow":
                lines.append("indices: {" + ", ".join(f"{q(a, next(qi))}: {q(b, next(qi))}"
                                                      for a, b in v.items()) + "}" + tail)
            else:
                lines.append("indices:" + tail)
                lines += [f"  {q(a, next(qi))}: {q(b, next(qi))}" for a, b in v.items()]
        elif key == "comment":
            lines.append(f'comment: "{v}"')
        elif key in ("start", "end"):
            lines.append(f"{key}: {q(v, next(qi))}{tail}")
        else:
            lines.append(f"{key}: {v!r}{tail}")
    return "\n".join(lines) + "\n"


@SETTINGS
@given(configs())
def test_config_loader_matches_yaml(mcr, tmp_path, cfg):
    c, style, quotes, trailing, order = cfg
    text = render(c, style, quotes, trailing, order)
    path = tmp_path / "config.yaml"
    path.write_text(text)
    got = mcr.load_config(path)
    # the text is YAML with exactly the keys written (the values are compared
    # with what was written: PyYAML would resolve unquoted `no` or `null` to
    # booleans and None, the reference's loader keeps scalars as text)
    ref = yaml.safe_load(text) or {}
    assert set(ref) == {k for k, v in c.items() if v is not None}
    assert got["models"] == c["models"]
    assert got["indices"] == c["indices"] and list(got["indices"]) == list(c["indices"])
    for key in ("start", "end"):
        assert got[key] == (None if c[key] is None else c[key])
    assert got["comment"] == c["comment"]
    # scalars as the reference converts them (atof / atoi / atol), defaults
    assert got["x"] == (0.02 if c["x"] is None else float(repr(c["x"])))
    assert got["train_ratio"] == (0.8 if c["train_ratio"] is None else float(repr(c["train_ratio"])))
    assert got["M"] == (100000 if c["M"] is None else c["M"])
    assert got["k"] == (5 if c["k"] is None else c["k"])
    assert got["N"] == (0 if c["N"] is None else c["N"])
    assert got["random_seed"] == (42 if c["random_seed"] is None else c["random_seed"])


row_text = st.text(alphabet=st.characters(min_codepoint=32, max_codepoint=126)
                   | st.sampled_from(["\n", "\r", '"', ","]), max_size=30)


@SETTINGS
@given(comment=row_text, names=st.lists(st.text(
    alphabet=st.characters(min_codepoint=33, max_codepoint=126, blacklist_characters=";:"),
    min_size=1, max_size=6), max_size=3, unique=True))
def test_csv_row_reads_back(mcr, comment, names):
    indices = {nm: f"{nm}.csv" for nm in names}
    line = mcr.csv_format_row(iteration_id=1, timestamp=1700000000, execution_time_ms=5,
                              comment=comment, start_date="2015-01-01", end_date="2025-12-31",
                              train_ratio=0.8, M=1000, k=5, x=0.02, model_name="hip", seed=7,
                              nodes=1, threads=1, processes=1, indices=indices,
                              actual_freq=0.01, P_hat=0.5, count=500, std_error=0.01,
                              ci_lower=0.4, ci_upper=0.6)
    rows = list(csv.reader(io.StringIO(line, newline="")))
    assert len(rows) == 1 and len(rows[0]) == 22
    assert rows[0][3] == comment
    assert rows[0][15] == ";".join(f"{k}:{v}" for k, v in indices.items())
    assert rows[0][10] == "hip" and rows[0][18] == "500"


yaml_ish = st.text(alphabet=st.sampled_from(list("abkxMN_:-[]{},'\"# \t\n0123456789.e")),
                   max_size=200)


@SETTINGS
@given(text=st.one_of(yaml_ish, st.binary(max_size=200).map(lambda b: b.decode("latin-1"))))
def test_config_loader_survives_any_text(mcr, tmp_path, text):
    """Arbitrary bytes and YAML-like fragments (unbalanced brackets, stray
    quotes, tabs, keys without values): the loader returns a configuration or
    MCR_E_ARG, never anything else (and, under the sanitizer run, touches no
    memory it does not own)."""
    path = tmp_path / "fuzz.yaml"
    path.write_bytes(text.encode("latin-1"))
    try:
        c = mcr.load_config(path)
    except mcr.McrError as e:
        assert e.code == mcr.MCR_E_ARG
    else:
        assert isinstance(c["models"], list) and isinstance(c["indices"], dict)


@st.composite
def bin_files(draw):
    """A header that may lie about the payload, then a payload that may be
    short, long or not numbers at all."""
    import struct

    rows = draw(st.integers(-3, 40))
    cols = draw(st.integers(-3, 40))
    kind = draw(st.sampled_from(["mu", "sigma", "params"]))
    head = struct.pack("<i", rows) if kind == "mu" else (
        struct.pack("<ii", rows, cols) if kind == "sigma" else b"")
    body = draw(st.binary(max_size=8 * 60))
    return kind, head + body


@SETTINGS
@given(bin_files())
def test_binary_readers_survive_any_file(mcr, tmp_path, case):
    """Headers that disagree with the file size, negative or zero counts, NaN
    and out-of-range values: each reader returns its array or MCR_E_ARG /
    MCR_E_IO-style errors, never a crash or a read past the file."""
    kind, data = case
    path = tmp_path / f"{kind}.bin"
    path.write_bytes(data)
    reader = {"mu": mcr.read_mu_bin, "sigma": mcr.read_sigma_bin,
              "params": mcr.read_params_bin}[kind]
    import math
    import struct

    try:
        out = reader(path)
    except mcr.McrError as e:
        assert e.code < 0
    else:
        # what loaded is what the header announced, read from bytes the file
        # has (trailing bytes are ignored, as the reference's fread does)
        if kind == "mu":
            (n,) = struct.unpack("<i", data[:4])
            assert out.shape == (n,) and len(data) >= 4 + 8 * n
        elif kind == "sigma":
            r, c = struct.unpack("<ii", data[:8])
            assert out.shape == (r, c) and len(data) >= 8 + 8 * r * c
        else:
            # the reference's range test lets a NaN through
            assert len(data) >= 8 and (math.isnan(out) or 0.0 <= out <= 1.0)
