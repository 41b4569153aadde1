#!/bin/sh
# Writes a configured copy of the reference's flag headers into $OUT.
# HARM is configured by editing these headers (global.h, metric.h, step_ch.h);
# only #define lines are changed, the .c files are compiled where they lie.
# With EXTRAP given, it also writes a copy of bounds.fishmon.c with only its
# "#define EXTRAP" line changed (the file defines it itself).
#   usage: configure.sh REFDIR OUT N1 N2 MCOORD FIXUPFLUX [EOMTYPE [CHECKSOLUTION [DOENO [EXTRAP]]]]
set -e
REF=$1; OUT=$2; N1=$3; N2=$4; MCOORD=$5; FIXUPFLUX=$6; EOM=${7:-EOMGRMHD}; CHECKSOL=${8:-0}; DOENO=${9:-0}; EXTRAP=${10:-}
mkdir -p "$OUT"
sed -e "s/^#define N1\([[:space:]]\).*/#define N1 $N1/" \
    -e "s/^#define N2\([[:space:]]\).*/#define N2 $N2/" \
    -e 's/^#define PRODUCTION 0/#define PRODUCTION 1/' \
    -e 's/^#define MAILWHENDONE 1/#define MAILWHENDONE 0/' \
    -e "s/^#define EOMTYPE EOMFFDE/#define EOMTYPE $EOM/" \
    -e 's/^#define UTOPRIMTRYAGAIN 1/#define UTOPRIMTRYAGAIN 0/' \
    -e "s/^#define CHECKSOLUTION 1/#define CHECKSOLUTION $CHECKSOL/" \
    -e "s/^#define DOENO 0/#define DOENO $DOENO/" \
    "$REF/global.h" > "$OUT/global.h"
sed -e "s/^#define MCOORD .*/#define MCOORD $MCOORD/" "$REF/metric.h" > "$OUT/metric.h"
sed -e "s/^#define FIXUPFLUX 1/#define FIXUPFLUX $FIXUPFLUX/" "$REF/step_ch.h" > "$OUT/step_ch.h"
# each edit must have matched exactly once
for pat in "define N1 $N1" "define N2 $N2" "define PRODUCTION 1" "define MAILWHENDONE 0" \
           "define EOMTYPE $EOM" "define UTOPRIMTRYAGAIN 0" "define CHECKSOLUTION $CHECKSOL" \
           "define DOENO $DOENO"; do
  test "$(grep -c "^#$pat\$" "$OUT/global.h")" = 1 || { echo "configure: '$pat' not applied"; exit 1; }
done
grep -q "^#define MCOORD $MCOORD" "$OUT/metric.h"
grep -q "^#define FIXUPFLUX $FIXUPFLUX" "$OUT/step_ch.h"
if [ -n "$EXTRAP" ]; then
  sed -e "s/^#define EXTRAP 0$/#define EXTRAP $EXTRAP/" "$REF/bounds.fishmon.c" > "$OUT/bounds.fishmon.c"
  test "$(grep -c "^#define EXTRAP $EXTRAP\$" "$OUT/bounds.fishmon.c")" = 1 || { echo "configure: EXTRAP not applied"; exit 1; }
fi
