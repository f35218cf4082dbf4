#!/bin/bash
# Aging workflow of run-fself.sh (reference run-fself.sh:1-197) on the MI355X
# engine, in one run: the configs are generated by bd_aging_self
# --make-configs (the reference calls three external binaries that it does not
# contain, run-fself.sh:131-138), and instead of one reheat-and-restart run per
# decade (run-fself.sh:149-197) a single run measures F_s and the MSD from every
# decade's waiting time (bd_aging_self --aging).  Output layout as the
# reference's: simulation/results/decadeK/{fself,msd_averaged}.out, plus
# decade 0's gder_k.out, energies_averaged.out and pressure.out, and the
# decade configuration dumps under configs/decadeK/.
#
#   usage: run_aging.sh SAMPLE_TAG     (run from a directory of your choice)
#          BD_AGING_PARAMS=file run_aging.sh SAMPLE_TAG   (parameters overridden)
set -e

#########################Parameters#########################
# configurations
ngrain_big=1024
ngrain_sml=1024
phi=0.58
sigma_big=1.4
sigma_sml=1.0
number_configs=5
eq_steps_configs=1000     # equilibration of each configuration at temp0
tagb=8
tags=8

# temperature
type="linear"
temp0=1.0
tempf=0.25
t0=0.0
tf=0.0
ngap_rescaling=1

# simulation
sampling="log"
dt_aging=1.0e-4
transtime_aging=0.1
runtime_aging=1.0e1
initial_decade=1.0e-3
number_initial_decades=3  # at most 8 waiting times in one run
samples_aging=10

# g(r) and F_s
gr_nbins=400
gr_range=15.0
qmax=6.19

# optional overrides: a file of name=value lines for the parameters above
# (a shorter run, another density, ...)
if [ -n "${BD_AGING_PARAMS:-}" ]; then
  # shellcheck disable=SC1090
  . "$BD_AGING_PARAMS"
fi

BIN="${BD_AGING_SELF:-$(cd "$(dirname "$0")" && pwd)/bd_aging_self}"

if [ $# -eq 0 ]; then
  echo "Error, no sample tag given"
  exit 1
fi
directorio=bigsml_${ngrain_big}_${ngrain_sml}_phi_${phi}_temp_${temp0}_${tempf}_${tf}_sample$1
if [ -d "$directorio" ]; then
  echo "Directory exists, use another tag"
  exit 2
fi
mkdir "$directorio"
cp "$0" "$directorio/"
cd "$directorio"

echo "Creating files"
for ((i = 0; i <= number_initial_decades; i++)); do
  mkdir -p "configs/decade$i" "simulation/results/decade$i"
done

echo "Creating all configs"
"$BIN" --make-configs "$ngrain_big" "$ngrain_sml" "$phi" "$sigma_big" "$sigma_sml" \
  "$number_configs" "$RANDOM" configs/decade0 "$temp0" "$eq_steps_configs" "$dt_aging"

echo "Starting simulation"
cd simulation
{
  echo "$type"
  echo "$temp0  $tempf  $t0  $tf  $ngap_rescaling"
  echo "$dt_aging"
  echo "$sampling"
  if [ "$sampling" = "linear" ]; then
    echo "$transtime_aging  $runtime_aging  $samples_aging"
  else
    echo "$transtime_aging  $initial_decade  $runtime_aging"
  fi
  echo "1  $number_initial_decades  0"
  echo "$tagb  $tags"
  echo "$number_configs"
  echo "$gr_nbins  $gr_range"
  echo "$qmax"
  echo "$RANDOM"
} > bd_aging_self.data

if ! "$BIN" --aging < bd_aging_self.data; then
  echo "Error in simulation"
  exit 4
fi
echo "Simulation finished successfully"
