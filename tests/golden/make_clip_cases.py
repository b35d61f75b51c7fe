"""Writes tests/golden/clip_cases.json from the reference clipper.

Reads src/shaders/polygon_clipping.glsl of the reference as text (it is
GLSL, nothing is executed) and records, for every case of its switch and for
every MAX_POLYGON_VERTEX_COUNT / MIN_POLYGON_VERTEX_COUNT_BEFORE_CLIPPING
variant, the input (vertex count, inside mask) and the output: the clipped
vertex count and, per output slot 0..count, which input vertex or which
edge intersection iz0(v[a], v[b]) it holds. The file is data (inputs and
expected outputs); the oracle's clipping rule is checked against it.

Usage: python tests/golden/make_clip_cases.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "src", "shaders", "polygon_clipping.glsl")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "clip_cases.json")


COMPARE = {"<=": lambda a, b: a <= b, ">=": lambda a, b: a >= b, "==": lambda a, b: a == b,
           "!=": lambda a, b: a != b, "<": lambda a, b: a < b, ">": lambda a, b: a > b}
TERM = re.compile(r"^\s*([A-Z_]+)\s*(<=|>=|==|!=|<|>)\s*(\d+)\s*$")


def evaluate_condition(cond, values):
    """An #if condition of the clipper: a conjunction (&&) of comparisons of
    a known macro with an integer literal. Anything else is rejected; the
    text is parsed, never executed."""
    result = True
    for term in cond.split("&&"):
        m = TERM.match(term)
        if not m or m.group(1) not in values:
            raise ValueError("unsupported #if condition: %r" % cond)
        result = result and COMPARE[m.group(2)](values[m.group(1)], int(m.group(3)))
    return result


def cases_for(text, min_count, max_count):
    out = {}
    stack = []
    values = {"MIN_POLYGON_VERTEX_COUNT_BEFORE_CLIPPING": min_count, "MAX_POLYGON_VERTEX_COUNT": max_count}

    def ev(cond):
        return evaluate_condition(cond, values)

    for line in text.split("\n"):
        s = line.strip()
        if s.startswith("#if "):
            v = ev(s[4:])
            stack.append([v, v])
        elif s.startswith("#elif "):
            top = stack[-1]
            if top[1]:
                top[0] = False
            else:
                v = ev(s[6:])
                top[0] = top[1] = v
        elif s.startswith("#endif"):
            stack.pop()
        elif s.startswith("case") and all(c[0] for c in stack):
            m = re.match(r"case\s+(\d+):\s+vc = (\d+);(.*)break;", s)
            mask, vc, body = int(m.group(1)), int(m.group(2)), m.group(3)
            slots = [["v", i] for i in range(max_count)]
            for st in [x.strip() for x in body.split(";") if x.strip()]:
                mm = re.match(r"v\[(\d+)\] = (.*)", st)
                dst, rhs = int(mm.group(1)), mm.group(2)
                mi = re.match(r"iz0\(v\[(\d+)\], v\[(\d+)\]\)", rhs)
                if mi:
                    a, b = slots[int(mi.group(1))], slots[int(mi.group(2))]
                    assert a[0] == "v" and b[0] == "v"
                    slots[dst] = ["iz0", a[1], b[1]]
                else:
                    slots[dst] = list(slots[int(re.match(r"v\[(\d+)\]", rhs).group(1))])
            n = mask & 7
            inside = mask >> 3
            keep = vc + 1 if vc < max_count else vc
            out["%d/%d" % (n, inside)] = {"n": n, "inside_mask": inside, "count": vc,
                                          "slots": slots[:keep] if vc else []}
    return out


def main():
    text = open(SRC).read()
    data = {}
    for max_count in range(4, 10):
        for min_count in range(3, max_count):
            data["max%d_min%d" % (max_count, min_count)] = cases_for(text, min_count, max_count)
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=0, sort_keys=True)
    print("wrote", OUT, sum(len(v) for v in data.values()), "cases")


if __name__ == "__main__":
    main()
