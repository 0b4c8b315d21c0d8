"""Golden fixtures: tiny cases in the reference's text formats with expected
outputs from complete enumeration (see make_golden.py)."""
