"""Pattern text -> Thompson NFA -> subset DFA -> complete, minimal DFA.

Host-side construction for the matching engine. It accepts the same pattern
syntax as the reference parser (``src/regex_to_dfa.py:99-210``):

* concatenation, ``|``, postfix ``*`` ``+`` ``?`` (one operator per atom; a
  second one is read as a literal character, as in the reference);
* ``(...)`` groups (an unclosed group is accepted; an unmatched ``)`` ends the
  pattern);
* ``[...]`` / ``[^...]`` classes with ``a-z`` ranges, no escapes inside a
  class; negation is taken over the 95 printable ASCII characters; an empty
  class matches ``"\\x00"``;
* ``.`` = the 95 printable ASCII characters; ``\\d \\w \\s``; ``\\x`` = literal x.

Differences from the reference, all of which leave the accepted language
unchanged:

* a character class is one NFA edge labelled with a set of characters instead
  of an alternation of one-character fragments, so NFAs are smaller;
* subset construction walks the alphabet in sorted order and numbers states in
  breadth-first order, so state ids do not depend on Python's hash seed
  (the reference iterates a ``set``, ``src/regex_to_dfa.py:255,263``);
* :func:`minimize_dfa` always minimizes, including when every state or no
  state accepts (the reference skips that case, ``src/regex_to_dfa.py:303-305``),
  and renumbers the result canonically (breadth-first from the start state);
* ``DFA.dead_state`` names the absorbing non-accepting state whenever the
  minimal DFA has one.

Matching batches of input bytes is done on the GPU by :mod:`trex_amd.engine`;
this module builds tables, and keeps the reference's single-string
``DFA.simulate`` as the construction type's own membership test.
"""

from __future__ import annotations

import string
from collections import deque
from dataclasses import dataclass, field
from typing import Optional

EPSILON = None

PRINTABLE95 = string.printable[:95]
_CLASS_D = "0123456789"
_CLASS_W = string.ascii_letters + string.digits + "_"
_CLASS_S = " \t\n\r"


# ---------------------------------------------------------------------------
# NFA
# ---------------------------------------------------------------------------

@dataclass
class NFAState:
    id: int
    transitions: dict = field(default_factory=dict)  # symbol or None -> list[int]
    is_accept: bool = False


class NFA:
    """Thompson NFA; ``transitions[None]`` holds the epsilon edges."""

    def __init__(self):
        self.states: dict[int, NFAState] = {}
        self.start: int = -1
        self.accept: int = -1
        self.alphabet: set[str] = set()

    def new_state(self, is_accept: bool = False) -> int:
        sid = len(self.states)
        self.states[sid] = NFAState(id=sid, is_accept=is_accept)
        return sid

    def add_trans(self, src: int, sym: Optional[str], dst: int) -> None:
        self.states[src].transitions.setdefault(sym, []).append(dst)
        if sym is not None:
            self.alphabet.add(sym)

    def add_set_trans(self, src: int, chars, dst: int) -> None:
        for ch in sorted(set(chars)):
            self.add_trans(src, ch, dst)


# ---------------------------------------------------------------------------
# Parser: pattern text -> AST
#   ('set', frozenset[str])   one character out of a set
#   ('eps',)                  empty string
#   ('cat', [ast, ...])
#   ('alt', [ast, ...])
#   ('star'|'plus'|'opt', ast)
# ---------------------------------------------------------------------------

class RegexParser:
    """Recursive-descent parser with the reference grammar
    (``src/regex_to_dfa.py:99-210``). ``parse()`` returns the Thompson NFA."""

    def __init__(self, pattern: str):
        self.pattern = pattern
        self.pos = 0
        self.nfa = NFA()

    # -- scanning --
    def _peek(self) -> Optional[str]:
        return self.pattern[self.pos] if self.pos < len(self.pattern) else None

    def _take(self) -> str:
        if self.pos >= len(self.pattern):
            raise ValueError(f"pattern {self.pattern!r} ends inside an escape")
        ch = self.pattern[self.pos]
        self.pos += 1
        return ch

    # -- grammar --
    def parse_ast(self):
        return self._alternation()

    def _alternation(self):
        branches = [self._sequence()]
        while self._peek() == "|":
            self.pos += 1
            branches.append(self._sequence())
        return branches[0] if len(branches) == 1 else ("alt", branches)

    def _sequence(self):
        items = []
        while self._peek() is not None and self._peek() not in "|)":
            items.append(self._postfix())
        if not items:
            return ("eps",)
        return items[0] if len(items) == 1 else ("cat", items)

    def _postfix(self):
        node = self._atom()
        op = self._peek()
        if op is not None and op in "*+?":
            self.pos += 1
            return ({"*": "star", "+": "plus", "?": "opt"}[op], node)
        return node

    def _atom(self):
        ch = self._take()
        if ch == "(":
            inner = self._alternation()
            if self._peek() == ")":
                self.pos += 1
            return inner
        if ch == "[":
            return ("set", frozenset(self._class_body()))
        if ch == ".":
            return ("set", frozenset(PRINTABLE95))
        if ch == "\\":
            esc = self._take()
            table = {"d": _CLASS_D, "w": _CLASS_W, "s": _CLASS_S}
            return ("set", frozenset(table.get(esc, esc)))
        return ("set", frozenset(ch))

    def _class_body(self) -> set[str]:
        negate = False
        if self._peek() == "^":
            negate = True
            self.pos += 1
        members: set[str] = set()
        while self._peek() is not None and self._peek() != "]":
            lo = self._take()
            nxt = self.pos + 1
            if self._peek() == "-" and nxt < len(self.pattern) and self.pattern[nxt] != "]":
                self.pos += 1
                hi = self._take()
                members.update(chr(c) for c in range(ord(lo), ord(hi) + 1))
            else:
                members.add(lo)
        if self._peek() == "]":
            self.pos += 1
        if negate:
            members = set(PRINTABLE95) - members
        return members or {"\x00"}

    # -- Thompson construction --
    def parse(self) -> NFA:
        ast = self.parse_ast()
        start, end = _thompson(self.nfa, ast)
        self.nfa.start = start
        self.nfa.accept = end
        self.nfa.states[end].is_accept = True
        return self.nfa


def _thompson(nfa: NFA, node) -> tuple[int, int]:
    kind = node[0]
    if kind == "set":
        s, e = nfa.new_state(), nfa.new_state()
        nfa.add_set_trans(s, node[1], e)
        return s, e
    if kind == "eps":
        s, e = nfa.new_state(), nfa.new_state()
        nfa.add_trans(s, EPSILON, e)
        return s, e
    if kind == "cat":
        frags = [_thompson(nfa, n) for n in node[1]]
        for (_, prev_end), (nxt_start, _) in zip(frags, frags[1:]):
            nfa.add_trans(prev_end, EPSILON, nxt_start)
        return frags[0][0], frags[-1][1]
    if kind == "alt":
        s, e = nfa.new_state(), nfa.new_state()
        for bs, be in (_thompson(nfa, n) for n in node[1]):
            nfa.add_trans(s, EPSILON, bs)
            nfa.add_trans(be, EPSILON, e)
        return s, e
    inner_s, inner_e = _thompson(nfa, node[1])
    s, e = nfa.new_state(), nfa.new_state()
    nfa.add_trans(s, EPSILON, inner_s)
    nfa.add_trans(inner_e, EPSILON, e)
    if kind in ("star", "opt"):
        nfa.add_trans(s, EPSILON, e)          # zero occurrences
    if kind in ("star", "plus"):
        nfa.add_trans(inner_e, EPSILON, inner_s)  # repeat
    return s, e


# ---------------------------------------------------------------------------
# DFA
# ---------------------------------------------------------------------------

@dataclass
class DFA:
    """Deterministic automaton over the characters in ``alphabet``.

    ``transitions[s][ch]`` is the successor of state ``s`` on character
    ``ch``. Characters outside ``alphabet`` lead to ``dead_state`` (an
    absorbing, non-accepting state); when ``dead_state`` is None the input is
    rejected, which is the same language.
    """

    n_states: int
    start: int
    accept_states: set[int]
    transitions: dict[int, dict[str, int]]
    alphabet: set[str]
    dead_state: Optional[int] = None

    def simulate(self, input_str: str) -> bool:
        """Whether the DFA accepts ``input_str``, walked one character at a
        time on the host (the reference's ``DFA.simulate``,
        ``src/regex_to_dfa.py:224-233``). A character with no transition moves
        to ``dead_state``; without a dead state it rejects at once. This is the
        construction type's own membership test for single strings; batches go
        to the GPU engines."""
        trans = self.transitions
        dead = self.dead_state
        state = self.start
        for ch in input_str:
            nxt = trans.get(state, {}).get(ch)
            if nxt is None:
                if dead is None:
                    return False
                nxt = dead
            state = nxt
        return state in self.accept_states


def epsilon_closure(nfa: NFA, states) -> frozenset[int]:
    seen = set(states)
    todo = list(seen)
    while todo:
        for nxt in nfa.states[todo.pop()].transitions.get(EPSILON, ()):
            if nxt not in seen:
                seen.add(nxt)
                todo.append(nxt)
    return frozenset(seen)


def nfa_to_dfa(nfa: NFA) -> DFA:
    """Subset construction; states are numbered in breadth-first order."""
    symbols = sorted(nfa.alphabet)
    first = epsilon_closure(nfa, [nfa.start])
    ids = {first: 0}
    order = [first]
    queue = deque([first])
    trans: dict[int, dict[str, int]] = {}
    while queue:
        cur = queue.popleft()
        row = trans.setdefault(ids[cur], {})
        for ch in symbols:
            moved = set()
            for s in cur:
                moved.update(nfa.states[s].transitions.get(ch, ()))
            if not moved:
                continue
            tgt = epsilon_closure(nfa, moved)
            if tgt not in ids:
                ids[tgt] = len(order)
                order.append(tgt)
                queue.append(tgt)
            row[ch] = ids[tgt]
    accept = {i for i, subset in enumerate(order) if nfa.accept in subset}
    return DFA(n_states=len(order), start=0, accept_states=accept,
               transitions=trans, alphabet=set(nfa.alphabet))


def complete_dfa(dfa: DFA) -> DFA:
    """Give every (state, character) pair a successor, adding a dead state
    when some pair has none (``src/regex_to_dfa.py:279-299``)."""
    missing = any(ch not in dfa.transitions.get(s, {})
                  for s in range(dfa.n_states) for ch in dfa.alphabet)
    if missing:
        dead = dfa.n_states
        dfa.n_states += 1
        for s in range(dfa.n_states):
            row = dfa.transitions.setdefault(s, {})
            for ch in dfa.alphabet:
                row.setdefault(ch, dead)
        dfa.dead_state = dead
    elif dfa.dead_state is None:
        dfa.dead_state = find_sink(dfa)
    return dfa


def find_sink(dfa: DFA) -> Optional[int]:
    """The absorbing non-accepting state of a complete DFA, if any."""
    for s in range(dfa.n_states):
        if s in dfa.accept_states:
            continue
        row = dfa.transitions.get(s, {})
        if all(row.get(ch) == s for ch in dfa.alphabet):
            return s
    return None


def minimize_dfa(dfa: DFA) -> DFA:
    """Minimal equivalent of a complete DFA (partition refinement on the
    states reachable from the start), renumbered breadth-first from the start
    state over the sorted alphabet."""
    symbols = sorted(dfa.alphabet)
    reach = {dfa.start}
    queue = deque([dfa.start])
    while queue:
        s = queue.popleft()
        for ch in symbols:
            t = dfa.transitions[s][ch]
            if t not in reach:
                reach.add(t)
                queue.append(t)
    states = sorted(reach)
    block = {s: int(s in dfa.accept_states) for s in states}
    while True:
        sig = {s: (block[s],) + tuple(block[dfa.transitions[s][ch]] for ch in symbols)
               for s in states}
        names: dict[tuple, int] = {}
        new_block = {s: names.setdefault(sig[s], len(names)) for s in states}
        if len(names) == len(set(block.values())):
            block = new_block
            break
        block = new_block
    # canonical renumbering: breadth-first over blocks
    rep = {}
    for s in states:
        rep.setdefault(block[s], s)
    new_id = {block[dfa.start]: 0}
    queue = deque([block[dfa.start]])
    trans: dict[int, dict[str, int]] = {}
    while queue:
        b = queue.popleft()
        row = trans.setdefault(new_id[b], {})
        for ch in symbols:
            tb = block[dfa.transitions[rep[b]][ch]]
            if tb not in new_id:
                new_id[tb] = len(new_id)
                queue.append(tb)
            row[ch] = new_id[tb]
    accept = {new_id[block[s]] for s in states if s in dfa.accept_states}
    out = DFA(n_states=len(new_id), start=0, accept_states=accept,
              transitions=trans, alphabet=set(dfa.alphabet))
    out.dead_state = find_sink(out)
    return out


def compile_regex(pattern: str, minimize: bool = True) -> DFA:
    """Pattern text -> complete (and by default minimal) DFA
    (``src/regex_to_dfa.py:356-364``)."""
    dfa = complete_dfa(nfa_to_dfa(RegexParser(pattern).parse()))
    if minimize:
        dfa = complete_dfa(minimize_dfa(dfa))
    return dfa
