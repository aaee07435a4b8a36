"""Writes tests/golden/registry.json: the reference's field registry (the
state arena's slots in order, with shape and count) and its snapshot field
order, read out of the reference's TypeScript source text. Nothing of the
reference is imported or run: the script tokenises the few expression
forms the `fields()` functions and the snapshot ORDER use.

    python tests/golden/make_registry.py [REFERENCE_ROOT]

Sources read (under src/model/climate/gcm/): stepper/index.ts:38-50 (the
domain order), each domain's fields() (dynamics/index.ts:27-65,
physics/{fixer,surface-flux,radiation,precipitation,surface,land,glacier,
vegetation,sea-ice,ocean,sea,carbon,storm}/index.ts, snapshot/index.ts:
14-23), the counts' constants (calendar/date/index.ts CLIMATOLOGY_MONTHS,
the domains' SOIL_LAYERS / STEP_VALUES) and fields/index.ts:18-128 (ORDER,
fieldCount).

Output: {"slots": {"5": [[name, shape, count], ...], "10": [...]},
"snapshot": {"5": [id, ...], "10": [...]}}: counts are resolved for 5 and
10 levels.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
GCM = "src/model/climate/gcm"


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def matching(text, i):
    """Index just past the bracket that closes text[i]."""
    pairs = {"(": ")", "[": "]", "{": "}"}
    stack = [pairs[text[i]]]
    j = i + 1
    quote = None
    while j < len(text):
        c = text[j]
        if quote:
            if c == "\\":
                j += 1
            elif c == quote:
                quote = None
        elif c in "\"'`":
            quote = c
        elif c in pairs:
            stack.append(pairs[c])
        elif c in ")]}":
            if c != stack.pop():
                raise ValueError("unbalanced brackets")
            if not stack:
                return j + 1
        j += 1
    raise ValueError("unclosed bracket")


def split_top(text, sep=","):
    """Split at top-level separators."""
    out, depth, cur, quote = [], 0, "", None
    for c in text:
        if quote:
            cur += c
            if c == quote:
                quote = None
            continue
        if c in "\"'`":
            quote = c
        elif c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
        if c == sep and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += c
    if cur.strip():
        out.append(cur.strip())
    return out


class Module:
    def __init__(self, ref, path):
        self.ref = ref
        self.path = path
        self.text = strip_comments(open(path).read())
        self.imports = {}
        for names, src in re.findall(r"import\s*\{([^}]*)\}\s*from\s*\"@/([^\"]+)\"", self.text):
            for n in names.split(","):
                n = n.strip()
                if n and not n.startswith("type "):
                    self.imports[n] = src

    def module(self, name):
        src = self.imports[name]
        base = os.path.join(self.ref, "src", src)
        path = base + "/index.ts" if os.path.isdir(base) else base + ".ts"
        return Module(self.ref, path)

    def const_expr(self, name, scope=None):
        """The expression text of `const name = ...` (the function's body
        first, then the file)."""
        for text in ([scope] if scope else []) + [self.text]:
            m = re.search(r"\bconst\s+" + name + r"\b(?:\s*:[^=]+)?\s*=\s*", text)
            if not m:
                continue
            i = m.end()
            j = i
            while j < len(text):
                if text[j] in "([{":
                    j = matching(text, j)
                    continue
                if text[j] == "\n":
                    rest = text[j:].lstrip()
                    if not rest.startswith("."):
                        break
                j += 1
            return text[i:j].strip()
        raise KeyError(name)

    def number(self, name):
        return int(self.const_expr(name))

    # ---- string-array expressions
    def names(self, expr, scope=None):
        expr = expr.strip()
        expr = re.sub(r"\s+as\s+const\s*$", "", expr)
        if expr.startswith("["):
            end = matching(expr, 0)
            out = []
            for item in split_top(expr[1:end - 1]):
                if item.startswith("..."):
                    out += self.names(item[3:], scope)
                elif item[0] in "\"'":
                    out.append(item[1:-1])
                else:
                    raise ValueError(f"array item {item!r}")
            rest = expr[end:]
        else:
            m = re.match(r"[A-Za-z_]\w*", expr)
            out = self.names(self.const_expr(m.group(0), scope), scope)
            rest = expr[m.end():]
        rest = re.sub(r"^\s*as\s+const", "", rest)
        while rest.strip():
            rest = rest.strip()
            m = re.match(r"\.concat\(", rest)
            if not m:
                raise ValueError(f"chain {rest[:40]!r}")
            end = matching(rest, m.end() - 1)
            out += self.names(rest[m.end():end - 1], scope)
            rest = rest[end:]
        return out

    # ---- field specs
    def fields(self):
        m = re.search(r"function\s+fields\s*\([^)]*\)\s*:[^{]*\{", self.text)
        body = self.text[m.end() - 1:matching(self.text, m.end() - 1)]
        ret = body[body.index("return") + len("return"):-1].strip()
        return self.specs(ret, body)

    def spec_of_map(self, arr, lam, scope):
        shape = re.search(r"shape:\s*\"(\w+)\"", lam).group(1)
        count = re.search(r"count:\s*([^,}]+?)\s*[,}]", lam).group(1).strip()
        return [(n, shape, count) for n in self.names(arr, scope)]

    def specs(self, expr, scope):
        expr = expr.strip().rstrip(";").strip()
        if expr.startswith("[") and matching(expr, 0) == len(expr):
            out = []
            for item in split_top(expr[1:-1]):
                out += self.item(item, scope)
            return out
        # ARRAY.map((name) => ({ ... }))
        i = expr.rindex(".map(")
        lam = expr[i + 5:matching(expr, i + 4) - 1]
        return self.spec_of_map(expr[:i], lam, scope)

    def item(self, item, scope):
        if item.startswith("{"):
            name = re.search(r"name:\s*\"(\w+)\"", item).group(1)
            shape = re.search(r"shape:\s*\"(\w+)\"", item).group(1)
            count = re.search(r"count:\s*(.+?)\s*,?\s*\}\s*$", item, re.S).group(1).strip()
            return [(name, shape, count)]
        assert item.startswith("..."), item
        body = item[3:].strip()
        m = re.match(r"(\w+)\.fields\(", body)
        if m:
            return [(n, s, (self.module(m.group(1)), c)) for n, s, c in self.module(m.group(1)).fields()]
        if body.startswith("group("):
            arg = body[body.index("{"):]
            names = re.search(r"names:\s*(\w+)", arg).group(1)
            shape = re.search(r"shape:\s*\"(\w+)\"", arg).group(1)
            count = re.search(r"count:\s*([^,}]+?)\s*[,}]", arg).group(1).strip()
            return [(n, shape, count) for n in self.names(names, scope)]
        return self.specs(body, scope)


def resolve_count(mod, count, levels, snapshot_fields):
    if isinstance(count, tuple):
        mod, count = count
        if isinstance(count, tuple):
            return resolve_count(count[0], count[1], levels, snapshot_fields)
    c = count.replace(" ", "")
    if c == "levels":
        return levels
    if c == "levels+1":
        return levels + 1
    if re.fullmatch(r"\d+", c):
        return int(c)
    if c == "CALENDAR.CLIMATOLOGY_MONTHS":
        date = Module(mod.ref, os.path.join(mod.ref, GCM, "calendar/date/index.ts"))
        return date.number("CLIMATOLOGY_MONTHS")
    if c.startswith("GCM_FIELDS.fieldCount("):
        return len(snapshot_fields)
    if re.fullmatch(r"[A-Z_]+", c):
        return mod.number(c)
    raise ValueError(f"count {count!r} in {mod.path}")


def snapshot_order(ref, layers):
    """fields/index.ts ORDER, cut to fieldCount(layers) (:416-418)."""
    mod = Module(ref, os.path.join(ref, GCM, "fields/index.ts"))
    expr = mod.const_expr("ORDER")
    out = []
    for item in split_top(expr[1:matching(expr, 0) - 1]):
        if item[0] == '"':
            out.append(item[1:-1])
        elif item.startswith("...STORM_MODES.map("):
            out += re.findall(r"mode:\s*\"(\w+)\"", mod.const_expr("STORM_MODES"))
        elif item.startswith("...LEVELS.flatMap("):
            levels = [int(x) for x in re.findall(r"\d+", mod.const_expr("LEVELS"))]
            templates = re.findall(r"`(\w+)\$\{level\}`", item)
            for level in levels:
                out += [t + str(level) for t in templates]
        else:
            raise ValueError(item)
    cut = re.search(r"return ORDER\.length - (\d+) \* \((\d+) - layers\)", mod.text)
    per, top = int(cut.group(1)), int(cut.group(2))
    return out[:len(out) - per * (top - layers)]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    stepper = Module(ref, os.path.join(ref, GCM, "stepper/index.ts"))
    specs = stepper.fields()
    out = {"source": "reference TypeScript text: stepper/index.ts fields() and the domains' fields(); "
                     "fields/index.ts ORDER",
           "slots": {}, "snapshot": {}}
    for levels in (5, 10):
        snap = snapshot_order(ref, levels)
        out["snapshot"][str(levels)] = snap
        out["slots"][str(levels)] = [[n, s, resolve_count(stepper, c, levels, snap)] for n, s, c in specs]
    path = os.path.join(HERE, "registry.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {path}: {len(specs)} slots, {len(out['snapshot']['10'])}/{len(out['snapshot']['5'])} "
          "snapshot fields")


if __name__ == "__main__":
    main()
