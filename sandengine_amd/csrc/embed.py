"""Embeds a source file as a C++ raw string literal (build step of the Makefile)."""
import sys

src, dst = sys.argv[1], sys.argv[2]
text = open(src).read()
delim = "SEKERNELS"
assert ")" + delim + '"' not in text
chunks = [text[i:i + 12000] for i in range(0, len(text), 12000)]
with open(dst, "w") as f:
    for c in chunks:
        f.write('R"' + delim + "(" + c + ")" + delim + '"\n')
