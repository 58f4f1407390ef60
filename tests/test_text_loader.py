"""The text-format circuit reader (Circuit::load_extracted_gates, reference
include/circuit/circuit.hpp:152-215) on the host: shapes follow _compute_nb_vars
(circuit.hpp:124-149), and corrupt files are refused with an error code -- a gate count
that would wrap a size computation, one larger than the file can hold, truncated gates and
out-of-range indices.  No device is needed (gkrx_circuit_parse_text)."""
import pytest

import gkrx


def _write(tmp_path, mul_lines, add_lines):
    mp, ap = tmp_path / "mul.txt", tmp_path / "add.txt"
    mp.write_text("\n".join(mul_lines) + "\n")
    ap.write_text("\n".join(add_lines) + "\n")
    return str(mp), str(ap)


def test_shape_of_a_generated_circuit(tmp_path):
    layers = gkrx.synth_circuit(7, 3, 6, 500, False)
    mp, ap = str(tmp_path / "m.txt"), str(tmp_path / "a.txt")
    gkrx.write_text(layers, mp, ap)
    ni, no = gkrx.parse_text_shape(mp, ap)
    assert ni == [6, 6, 6] and no == [6, 6, 6]


def test_shape_follows_the_largest_index(tmp_path):
    # output layer first.  Layer 1: largest input id 5 -> next_pow_of_2 = 8 -> 3 variables,
    # largest output id 3 -> 4 -> 2 variables; layer 0: inputs up to 3 (2 variables),
    # outputs up to 7 (3 variables) (circuit.hpp:124-149, myutil.hpp:38-48)
    mp, ap = _write(tmp_path, ["1 5 0 3 1", "2 0 1 3 1 1 2 7 1"], ["0", "1 3 0 1"])
    ni, no = gkrx.parse_text_shape(mp, ap)
    assert ni == [2, 3] and no == [3, 2]


@pytest.mark.parametrize("count", [
    (1 << 62) + 1,            # 4 * count wraps to 4 in 64 bits
    (1 << 32),                # above the 32-bit range of a gate count
    (1 << 31),                # large, no wrap: more gates than the file holds
    1000,                     # more gates than the file holds
])
def test_corrupt_gate_count_is_refused(tmp_path, count):
    mp, ap = _write(tmp_path, [f"{count} 0 1 2 1 3 2 1 1"], ["0"])
    with pytest.raises(gkrx.GkrxError, match="gate count"):
        gkrx.parse_text_shape(mp, ap)
    mp, ap = _write(tmp_path, ["1 0 1 1 1"], [f"{count} 0 1 1"])
    with pytest.raises(gkrx.GkrxError, match="gate count"):
        gkrx.parse_text_shape(mp, ap)


@pytest.mark.parametrize("mul,add", [
    (["2 0 1 1 1 1"], ["0"]),          # second gate truncated
    (["1 0 1 x 1"], ["0"]),            # not a number
    (["1 0 1 4294967296 1"], ["0"]),   # index above 2^32
    (["1 0 1 1 1"], ["2 0 1 1"]),      # add file truncated
    (["1 0 1 1 1", "1 0 1 1 1"], ["0"]),  # add file has fewer lines than the mul file
])
def test_malformed_gates_are_refused(tmp_path, mul, add):
    mp, ap = _write(tmp_path, mul, add)
    with pytest.raises(gkrx.GkrxError):
        gkrx.parse_text_shape(mp, ap)


def test_coefficient_that_does_not_fold_is_refused(tmp_path):
    # M31(0xffffffff) folds to 2^31 (include/field/M31.hpp:58-62): refused for M31
    mp, ap = _write(tmp_path, ["1 0 3 3 4294967295"], ["0"])
    with pytest.raises(gkrx.GkrxError):
        gkrx.parse_text_shape(mp, ap)
    assert gkrx.parse_text_shape(mp, ap, gkrx.FIELD_BN254) == ([2], [2])


def test_missing_file(tmp_path):
    with pytest.raises(gkrx.GkrxError, match="cannot open"):
        gkrx.parse_text_shape(str(tmp_path / "none.txt"), str(tmp_path / "none2.txt"))
