"""The Java binding sources (finmath-lib_amd/java): no JDK exists in this image, so they are
checked here without compiling them.

- The byte offsets the Panama FFM binding writes (FinmathHip.java) are the C compiler's offsetof
  of fm_lmm_config and fm_product (include/finmath_lmm.h), compiled here with gcc.
- The binding reads the model only through methods the reference's
  LIBORMarketModelFromCovarianceModel declares public
  (src/main/java/net/finmath/montecarlo/interestrate/models/LIBORMarketModelFromCovarianceModel.java:
  844-1732); its state space and liborCap have no getters (private fields set from the properties
  map, :187-194, :400-414), so the binding probes applyStateSpaceTransform instead.
"""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JAVA = os.path.join(REPO, "finmath-lib_amd", "java", "net", "finmath", "montecarlo", "interestrate", "hip")

CONFIG_FIELDS = ["num_time_steps", "times", "num_libors", "tenor", "num_factors", "num_paths", "seed", "measure",
                 "state_space", "scheme", "precision", "libor_cap", "store_paths", "checkpoint_stride",
                 "regenerate_increments", "path_offset"]
PRODUCT_FIELDS = ["type", "num_periods", "exercise_date", "fixing_dates", "payment_dates", "period_lengths",
                  "swaprates", "notionals", "discounting_adjustment", "notional", "is_pay_fix", "is_floorlet",
                  "exercise_flags", "is_cancelable", "basis"]

# public methods of LIBORMarketModelFromCovarianceModel (and its ProcessModel / LIBORMarketModel
# interfaces) that take no process-internal state: the only ones the binding may call on the model
MODEL_PUBLIC = {"getReferenceDate", "getNumeraire", "getForwardDiscountBond", "getNumeraireAdjustments",
                "getInitialState", "getDrift", "getFactorLoading", "applyStateSpaceTransform",
                "applyStateSpaceTransformInverse", "getRandomVariableForConstant", "getDriftApproximationMethod",
                "getForwardRate", "getLIBOR", "getNumberOfComponents", "getNumberOfLibors", "getNumberOfFactors",
                "getLiborPeriod", "getLiborPeriodIndex", "getLiborPeriodDiscretization", "getInterpolationMethod",
                "getMeasure", "getIntegratedLIBORCovariance", "getAnalyticModel", "getDiscountCurve",
                "getForwardRateCurve", "getSwaptionMarketData", "getCovarianceModel",
                "getCloneWithModifiedCovarianceModel", "getCloneWithModifiedData", "getModelParameters"}


@pytest.fixture(scope="module")
def c_offsets(tmp_path_factory):
    d = tmp_path_factory.mktemp("offsets")
    src = d / "offsets.c"
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "finmath_lmm.h"', "int main(void) {",
             '  printf("CONFIG_SIZE %zu\\n", sizeof(fm_lmm_config));',
             '  printf("PRODUCT_SIZE %zu\\n", sizeof(fm_product));']
    lines += [f'  printf("OFF_{f.upper()} %zu\\n", offsetof(fm_lmm_config, {f}));' for f in CONFIG_FIELDS]
    lines += [f'  printf("P_{f.upper()} %zu\\n", offsetof(fm_product, {f}));' for f in PRODUCT_FIELDS]
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = d / "offsets"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def java_constants():
    text = open(os.path.join(JAVA, "FinmathHip.java")).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"public static final long (\w+) = (\d+);", text)}


def test_struct_offsets_of_the_panama_binding_match_the_c_header(c_offsets):
    java = java_constants()
    for name, off in c_offsets.items():
        assert name in java, f"FinmathHip.java lacks {name}"
        assert java[name] == off, f"{name}: Java {java[name]}, C {off}"


def test_status_and_enum_constants_match_the_c_header():
    header = open(os.path.join(REPO, "include", "finmath_lmm.h")).read()
    c = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define (FM_\w+)\s+(\d+)", header)}
    text = open(os.path.join(JAVA, "FinmathHip.java")).read()
    java = {m.group(1): int(m.group(2)) for m in re.finditer(r"public static final int (FM_\w+) = (\d+);", text)}
    assert java, "no FM_ constants found"
    for k, v in java.items():
        assert c[k] == v, k


def test_binding_reads_the_model_through_public_methods_only():
    calls = set()
    for f in os.listdir(JAVA):
        text = open(os.path.join(JAVA, f)).read()
        calls |= set(re.findall(r"\bmodel\.(\w+)\(", text))
        calls |= set(re.findall(r"liborModel\(\)\.(\w+)\(", text))
    assert calls, "no model calls found"
    assert calls <= MODEL_PUBLIC, sorted(calls - MODEL_PUBLIC)
    for f in os.listdir(JAVA):
        text = open(os.path.join(JAVA, f)).read()
        assert "getStateSpace(" not in text and "getLiborCap(" not in text


def test_jni_natives_match_the_java_declarations():
    java = open(os.path.join(JAVA, "FinmathHipJNI.java")).read()
    natives = set(re.findall(r"public static native \w+ (\w+)\(", java))
    c = open(os.path.join(REPO, "finmath-lib_amd", "java", "jni", "finmath_lmm_jni.c")).read()
    defined = set(re.findall(r"JNI_NAME\((\w+)\)\(", c))
    assert natives and natives == defined
