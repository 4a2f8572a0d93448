"""Reader of OpenFOAM dictionary files (ASCII), as uniGasFoam's cases hold
them: constant/uniGasProperties, system/controlDict, boundariesDict,
uniGasInitialisationDict, hybridDecompositionDict, fieldPropertiesDict and
blockMeshDict.

The syntax covered is the one those files use: ``//`` and ``/* */``
comments, the FoamFile header, ``key value;`` entries whose value is one or
more tokens, sub-dictionaries ``key { ... }``, lists ``( ... )`` (optionally
count-prefixed, nested, or holding ``word { ... }`` entries such as the
``boundary { ... }`` items of boundariesDict), ``[ ... ]`` dimension sets,
quoted strings, and ``$name`` references to entries of the same file.
``#include``-style directives are refused (DictError).

Values come back as Python objects: numbers as int or float, words as str,
lists as Python lists, dictionaries as ``FoamDict`` (an ordered dict that
remembers every entry), a list item ``word { ... }`` as ``(word, FoamDict)``,
and an entry of several tokens (``hex (0 1 2 3 4 5 6 7) (8 8 8)
simpleGrading (1 1 1)``) as a ``Tokens`` list.
"""
from __future__ import annotations

import re


class DictError(ValueError):
    """A file that is not a dictionary this reader understands."""


class Tokens(list):
    """An entry value made of more than one token."""


class FoamDict(dict):
    """Ordered dictionary of one OpenFOAM dictionary level."""

    def sub(self, key):
        v = self.get(key)
        if not isinstance(v, FoamDict):
            raise DictError(f"sub-dictionary {key!r} missing")
        return v

    def sub_or_empty(self, key):
        v = self.get(key)
        return v if isinstance(v, FoamDict) else FoamDict()

    def scalar(self, key, default=None):
        if key not in self:
            if default is None:
                raise DictError(f"entry {key!r} missing")
            return default
        v = self[key]
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise DictError(f"entry {key!r} is not a number: {v!r}")
        return float(v)

    def label(self, key, default=None):
        return int(self.scalar(key, default))

    def word(self, key, default=None):
        if key not in self:
            if default is None:
                raise DictError(f"entry {key!r} missing")
            return default
        v = self[key]
        if not isinstance(v, str):
            raise DictError(f"entry {key!r} is not a word: {v!r}")
        return v

    def switch(self, key, default=None):
        if key not in self:
            if default is None:
                raise DictError(f"switch {key!r} missing")
            return default
        return as_switch(self[key], key)

    def vector(self, key, default=None):
        if key not in self:
            if default is None:
                raise DictError(f"entry {key!r} missing")
            return tuple(default)
        v = self[key]
        if not (isinstance(v, list) and len(v) == 3 and all(isinstance(x, (int, float)) for x in v)):
            raise DictError(f"entry {key!r} is not a vector: {v!r}")
        return tuple(float(x) for x in v)


# OpenFOAM Switch words (Switch.C)
_TRUE = {"yes", "on", "true", "y", "t", "any"}
_FALSE = {"no", "off", "false", "n", "f", "none"}


def as_switch(v, what="switch"):
    if isinstance(v, bool):
        return v
    if isinstance(v, (int, float)):
        return v != 0
    if isinstance(v, str):
        s = v.lower()
        if s in _TRUE:
            return True
        if s in _FALSE:
            return False
    raise DictError(f"{what}: {v!r} is not a switch")


_TOKEN = re.compile(r'"(?:[^"\\]|\\.)*"|[{}()\[\];]|[^\s{}()\[\];"]+')


def _strip_comments(text):
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            if j < 0:
                raise DictError("unterminated /* comment")
            out.append(" ")
            i = j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


def _atom(tok):
    if tok.startswith('"'):
        return tok[1:-1]
    try:
        if re.fullmatch(r"[+-]?\d+", tok):
            return int(tok)
        return float(tok)
    except ValueError:
        return tok


class _Parser:
    def __init__(self, text, path, counted=True):
        self.toks = _TOKEN.findall(_strip_comments(text))
        self.i = 0
        self.path = path
        self.top = None
        self.counted = counted

    def err(self, msg):
        return DictError(f"{self.path}: {msg} (token {self.i})")

    def peek(self, k=0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else None

    def take(self):
        t = self.peek()
        if t is None:
            raise self.err("unexpected end of file")
        self.i += 1
        return t

    def expect(self, t):
        got = self.take()
        if got != t:
            raise self.err(f"expected {t!r}, found {got!r}")

    def resolve(self, tok):
        if tok.startswith("$"):
            name = tok[1:].strip("{}")
            for d in (self.cur, self.top):
                if d is not None and name in d:
                    return d[name]
            raise self.err(f"unknown reference {tok}")
        return _atom(tok)

    def value_item(self):
        t = self.peek()
        if t == "(":
            return self.lst()
        if t == "[":
            self.take()
            items = []
            while self.peek() != "]":
                items.append(_atom(self.take()))
            self.take()
            return items
        if t == "{":
            self.take()
            return self.dict_body("}")
        tok = self.take()
        if tok in (")", "}", "]", ";"):
            raise self.err(f"unexpected {tok!r}")
        if tok.startswith("#"):
            raise self.err(f"directive {tok} is not supported")
        # a count-prefixed list: 3 ( ... ); a number before a list of another
        # length is a value of its own.  Files without counted lists (a
        # blockMeshDict, whose "arc 2 3 (x y z)" would read as one) turn
        # the detection off.
        if self.counted and self.peek() == "(" and re.fullmatch(r"\d+", tok) is not None:
            at = self.i
            items = self.lst()
            if len(items) == int(tok):
                return items
            self.i = at
        return self.resolve(tok)

    def lst(self):
        self.expect("(")
        items = []
        while self.peek() != ")":
            if self.peek() is None:
                raise self.err("unterminated list")
            t = self.peek()
            if t not in ("(", "{", "[") and self.peek(1) == "{" and not t.startswith('"'):
                name = self.take()
                self.take()
                items.append((name, self.dict_body("}")))
            else:
                items.append(self.value_item())
        self.take()
        return items

    def dict_body(self, end):
        d = FoamDict()
        saved = getattr(self, "cur", None)
        self.cur = d
        if self.top is None:
            self.top = d
        while True:
            t = self.peek()
            if t is None:
                if end is None:
                    break
                raise self.err("unterminated dictionary")
            if t == end:
                self.take()
                break
            if t == ";":
                self.take()
                continue
            key = self.take()
            if key.startswith("#"):
                raise self.err(f"directive {key} is not supported")
            if key in ("{", "}", "(", ")", "[", "]"):
                raise self.err(f"unexpected {key!r}")
            key = _atom(key) if key.startswith('"') else key
            if self.peek() == "{":
                self.take()
                d[key] = self.dict_body("}")
                continue
            vals = []
            while self.peek() != ";":
                if self.peek() is None or self.peek() == end:
                    raise self.err(f"entry {key!r} not terminated by ';'")
                vals.append(self.value_item())
            self.take()
            d[key] = vals[0] if len(vals) == 1 else Tokens(vals)
        self.cur = saved
        return d


def parse(text, path="<string>", counted=True) -> FoamDict:
    """A dictionary from its text; counted=False reads "N ( ... )" as a
    number followed by a list (files that hold no counted lists)."""
    p = _Parser(text, path, counted)
    d = p.dict_body(None)
    d.pop("FoamFile", None)
    return d


def read(path, counted=True) -> FoamDict:
    with open(path) as f:
        text = f.read()
    return parse(text, path, counted)
