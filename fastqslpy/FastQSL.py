"""fastqslpy.FastQSL (fastqslpy/FastQSL.py:4-163 of peijin94/FastQSL):
CookPseudoLine, ResReshape, QCalcPlane, curl_of_B_grid, copy_boundary and
diff_3point with the reference's positional signatures, evaluated on the GPU
(see ``fastqsl_amd.FastQSL``)."""
from fastqsl_amd.FastQSL import (CookPseudoLine, QCalcPlane, ResReshape, copy_boundary, curl_of_B_grid,
                                 diff_3point)

__all__ = ["CookPseudoLine", "ResReshape", "QCalcPlane", "curl_of_B_grid", "copy_boundary", "diff_3point"]
