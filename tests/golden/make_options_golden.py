"""Write the option-format fixtures of tests/golden/ from the reference's own
options_utils.F90.

Run in the build container, where the reference checkout exists and
`make -C oracle ref` has built oracle/_ref/ref_options_driver (the
reference's options_utils module behind the driver oracle/ref_options_driver.f90):

    python tests/golden/make_options_golden.py

Fixtures (plain text):
  options_battery.opt      an options file of edge cases (input)
  options_battery_ref.txt  what load_options and the typed getters of the
                           reference read from it (the driver's "load" output)
  options_put.txt          typed put requests (input: kind name value)
  options_put_ref.opt      the file the reference's put_*_option and
                           save_options wrote for them
tests/test_io.py compares the library against these.
"""
import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_options_driver")


def battery_lines():
    """Comments, blank lines, spacing, truncation (80-character lines,
    20-character names, 60-character values), tabs, duplicates and the
    value spellings of integer, real and logical list-directed input."""
    return [
        "# comment", "   # indented comment", "", "    ",
        "s_plain value", "s_multi    spaced   value  ", "   s_lead leading blanks on the line",
        "i_int 42", "i_neg -17", "i_plus +5", "i_lead        7", "i_trail 8 extra words", "i_comma 9,10",
        "i_big 2147483647", "i_min -2147483647",
        "d_int 3", "d_dec 3.45", "d_exp 1e-4", "d_Dexp 1.5D+02", "d_dexp 2.5d-3", "d_noexp 1.0+5", "d_q 4.0Q1",
        "d_dot .5", "d_trail 5.", "d_neg -0.0", "d_F20 0.0039062500", "d_big 1.7976931348623157E+308",
        "d_sub 4.9E-324", "d_F20b         7.8900000000", "d_comma 1.25,7", "d_tiny 2.4703282292062328E-324",
        "d_half 0.1000000000000000055511151231257827", "d_many 3.14159265358979323846264338327950288",
        "l_T T", "l_true true", "l_dotf .false.", "l_f f", "l_tpls tpls", "l_F F", "l_dott .T",
        "s_hash value # not a comment", "s_dup first", "s_dup second",
        "this_name_is_longer_than_twenty_chars value",
        "s_long " + "v" * 70,
        "s_line " + "w" * 90,
        "s_tab\tvalue after tab",
        "x" * 85,
        "s_last done",
    ]


def bits(d: float) -> str:
    return "%016X" % struct.unpack("<Q", struct.pack("<d", d))[0]


def put_requests():
    """I10 fields (sign, overflow to asterisks), F20.10 fields (rounding at
    the tenth decimal, negative zero, overflow, Inf and NaN), L1, strings and
    a replaced name."""
    rng = np.random.default_rng(3)
    req = [("i", "i_a", 2), ("i", "i_b", -2147483647), ("i", "i_c", 2147483647), ("i", "i_d", -2147483648),
           ("i", "i_e", 0), ("i", "i_f", -999999999), ("i", "i_g", -1000000000),
           ("d", "d_a", 3.45), ("d", "d_b", 7.89), ("d", "d_c", -0.0), ("d", "d_d", 1e-11), ("d", "d_e", -1e-11),
           ("d", "d_f", 5e-11), ("d", "d_g", 1.5e-10), ("d", "d_h", 2.5e-10), ("d", "d_i", 123456789.0),
           ("d", "d_j", 1234567890.0), ("d", "d_k", -123456789.0), ("d", "d_l", 1e300),
           ("d", "d_m", float("inf")), ("d", "d_n", float("-inf")), ("d", "d_o", float("nan")),
           ("d", "d_p", 1 / 3), ("d", "d_q", 0.00390625), ("d", "d_r", 1 / 96), ("d", "d_s", 0.5),
           ("d", "d_t", -0.5), ("d", "d_u", 0.123456789049999999), ("d", "d_v", 2.00000000005)]
    req += [("d", f"d_r{i}", float(x)) for i, x in enumerate(rng.standard_normal(40) * 10.0 ** rng.integers(-12, 9, 40))]
    req += [("l", "l_a", True), ("l", "l_b", False), ("s", "s_a", "blah"), ("s", "s_b", "two words"),
            ("i", "i_a", 5)]
    lines = []
    for k, n, v in req:
        if k == "d":
            lines.append(f"d {n} {bits(v)}")
        elif k == "l":
            lines.append(f"l {n} {'T' if v else 'F'}")
        else:
            lines.append(f"{k} {n} {v}")
    return lines


def main():
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} is not built: run `make -C oracle ref` where the reference checkout exists")
    bat = os.path.join(HERE, "options_battery.opt")
    with open(bat, "w") as f:
        f.write("\n".join(battery_lines()) + "\n")
    out = subprocess.run([DRIVER, "load", bat], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "options_battery_ref.txt"), "w") as f:
        f.write(out)
    puts = "\n".join(put_requests()) + "\n"
    with open(os.path.join(HERE, "options_put.txt"), "w") as f:
        f.write(puts)
    subprocess.run([DRIVER, "put", os.path.join(HERE, "options_put_ref.opt")], input=puts, text=True, check=True)
    print("wrote options_battery.opt, options_battery_ref.txt, options_put.txt, options_put_ref.opt")


if __name__ == "__main__":
    main()
