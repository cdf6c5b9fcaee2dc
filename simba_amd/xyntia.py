"""Instance JSON -> Xyntia trace JSON (scripts/convert_to_xyntia.py:7-91).

The first example becomes the "initial" trace and the others the "sampling"
traces, keyed by their position.  Each input i sits at location "mem<i>", the
outputs at EAX, EBX, ECX, EDX, then "out<i>"; sizes and values are written
as Python hex strings, values masked to their bit size.  Files are written
with a 4-space indent, one output file per input file of the same name.

    python -m simba_amd.xyntia <instances_dir> <out_dir>
"""
import argparse
import json
import sys
from pathlib import Path

_OUT_REGS = ("EAX", "EBX", "ECX", "EDX")


def _slot(location, item):
    bits = int(item["size"])
    return {"location": location, "size": hex(bits),
            "value": hex(int(item["value"]) & ((1 << bits) - 1))}


def trace(inputs, outputs):
    """One input/output pair (convert_to_xyntia.py:7-36)."""
    return {
        "inputs": {str(i): _slot("mem%d" % i, x) for i, x in enumerate(inputs)},
        "outputs": {str(i): _slot(_OUT_REGS[i] if i < len(_OUT_REGS) else "out%d" % i, y)
                    for i, y in enumerate(outputs)},
    }


def convert(instance):
    """Xyntia document of one instance dict (convert_to_xyntia.py:39-67)."""
    ex = instance.get("examples") or []
    if not ex:
        raise ValueError("no examples in the instance")
    return {"initial": trace(ex[0]["inputs"], ex[0]["output"]),
            "sampling": {str(k): trace(e["inputs"], e["output"]) for k, e in enumerate(ex[1:])}}


def convert_file(src, dst):
    with open(src) as f:
        doc = convert(json.load(f))
    with open(dst, "w") as f:
        json.dump(doc, f, indent=4)


def convert_folder(src_dir, dst_dir):
    """Every *.json of src_dir into dst_dir (convert_to_xyntia.py:70-74)."""
    src_dir, dst_dir = Path(src_dir), Path(dst_dir)
    dst_dir.mkdir(parents=True, exist_ok=True)
    done = []
    for p in sorted(src_dir.glob("*.json")):
        convert_file(p, dst_dir / p.name)
        print("Successfully converted %s to Xyntia format at %s" % (p, dst_dir / p.name))
        done.append(p.name)
    return done


def main(argv=None):
    ap = argparse.ArgumentParser(description="Convert MBA instance JSON to Xyntia JSON.")
    ap.add_argument("input", type=Path, help="directory of instance JSON files")
    ap.add_argument("output", type=Path, help="destination directory")
    a = ap.parse_args(argv)
    try:
        convert_folder(a.input, a.output)
    except (ValueError, KeyError) as e:
        print("Error: %s" % e)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
