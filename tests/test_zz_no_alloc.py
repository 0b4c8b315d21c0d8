"""The allocation-free parse entry point (csm_parse_text_device_static)
against csm_parse_text_device on the same text in HBM: the same records, on
two calls in a row (the second with a smaller input, so that stale contents
of the shared scratch array would show), plus the size limit and an error
line."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _text(rng, n):
    lines = []
    for _ in range(n):
        if rng.random() < 0.2:
            lines.append(f"{'v' if rng.random() < 0.5 else '-v'} {rng.integers(0, 1 << 20)} "
                         f"{rng.integers(0, 64)}")
        else:
            tok = "e" if rng.random() < 0.7 else "-e"
            lines.append(f"{tok} {rng.integers(0, 1 << 30)} {rng.integers(0, 1 << 30)} "
                         f"{rng.integers(0, 4)}")
    return ("\n".join(lines) + "\n").encode()


def _engine():
    from paracosm_amd.engine import Matching
    from tests.helpers import query_path2
    vl = np.zeros(4, np.uint32)
    m = Matching("parallel_graphflow", query_path2(), vl, np.asarray([0, 1], np.uint32),
                 np.asarray([1, 2], np.uint32), np.zeros(2, np.uint32))
    m.Preprocessing()
    return m


def _parse(m, text, static, cap=None):
    from paracosm_amd.engine import PARSE_STREAM, UPDATE_DTYPE
    n_lines = text.count(b"\n") + 1
    cap = n_lines if cap is None else cap
    d_text = m.device_buffer(max(len(text), 16))
    d_text.upload(np.frombuffer(text, np.uint8))
    d_out = m.device_buffer(16 * cap)
    n = m.parse_text_device(d_text.ptr, len(text), PARSE_STREAM, d_out.ptr, cap, static=static)
    out = np.zeros(n, UPDATE_DTYPE)
    if n:
        d_out.download(out.view(np.uint8))
    d_text.close()
    d_out.close()
    return out


def test_static_parse_matches_allocating_parse_twice():
    rng = np.random.default_rng(4)
    big, small = _text(rng, 200_000), _text(rng, 3_000)
    print("no-alloc parse: texts of", len(big), "and", len(small), "bytes", flush=True)
    with _engine() as m:
        for text in (big, small):
            want = _parse(m, text, static=False)
            got = _parse(m, text, static=True)
            assert len(got) == len(want) > 0
            np.testing.assert_array_equal(got, want)


def test_static_parse_limits_and_errors():
    from paracosm_amd.engine import CSM_E_CAPACITY, CsmError, ParseError
    with _engine() as m:
        # a malformed line is reported as by the allocating parse
        bad = b"e 1 2 3\ne 1 2\n"
        with pytest.raises(ParseError) as ei:
            _parse(m, bad, static=True)
        assert ei.value.line == 2
        # the next call starts from a clean error state
        assert len(_parse(m, b"e 1 2 3\n-e 4 5 6\n", static=True)) == 2
        # more than 262,144 lines: refused before parsing
        many = b"e 1 2 3\n" * 300_000
        with pytest.raises(CsmError) as ei:
            _parse(m, many, static=True)
        assert ei.value.code == CSM_E_CAPACITY
