"""Writes tests/golden/gl16.json: the 16-point Gauss-Legendre nodes and weights
as the reference's table lists them (Gauss_Legendre_Nodes_and_Weights.hpp,
case 16, lines 471-523), as decimal strings.  Data only; run in a container
that has the reference mounted."""
import json
import re
import sys

src = open(sys.argv[1] if len(sys.argv) > 1 else
           "/root/reference/3D/codes/Gauss_Legendre_Nodes_and_Weights.hpp").read().splitlines()[470:524]
nodes, weights = {}, {}
for line in src:
    m = re.match(r"\s*(nodes|weights)\[(\d+)\]\s*=\s*([-0-9.eE+]+);", line)
    if m:
        (nodes if m.group(1) == "nodes" else weights)[int(m.group(2))] = m.group(3)
json.dump({"source": "Gauss_Legendre_Nodes_and_Weights.hpp:471-523 (case 16)",
           "nodes": [nodes[i] for i in range(16)], "weights": [weights[i] for i in range(16)]},
          open(__file__.replace("make_gl16.py", "gl16.json"), "w"), indent=1)
