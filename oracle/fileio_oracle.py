"""Restatement of the reference's text readers — TEST INFRASTRUCTURE ONLY.

``read_lines_reference`` follows utils/fileio.py:55-64 (read_vbr) and
:66-93 (read_vbrc): each line is split at "=", the brackets and newline are
cut off with ``[1:-2]``, and every comma-separated token is converted with
Python's float() or int() into a list.  It is the CPU baseline of the text
I/O measurement and the checker of the native reader in tests.

Pinned by: it is the reference's own parse (Python's correctly rounded
float() and int()); tests/test_fileio_native.py checks the native reader
against it token by token and against the golden ``example.vbr`` /
``example.vbrc`` files.
"""
from __future__ import annotations


def parse_line_reference(line: str, conv):
    """One line as the reference parses it (empty body -> empty list, as in
    its read_vbrc)."""
    body = line.split("=")[1][1:-2]
    return list(map(conv, body.split(","))) if body != "" else []


def read_lines_reference(path, convs):
    """The first len(convs) lines of a file, parsed as the reference does."""
    with open(path, "r") as f:
        return [parse_line_reference(f.readline(), c) for c in convs]
