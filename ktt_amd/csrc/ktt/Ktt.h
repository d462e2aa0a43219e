/** @file Ktt.h
  * Umbrella header of the ktt API (Source/Ktt.h). Include this file only.
  */
#pragma once

#define KTT_VERSION_MAJOR 2
#define KTT_VERSION_MINOR 1
#define KTT_VERSION_PATCH 0
#define KTT_VERSION_STRING "2.1.0-hip"

#include <ComputeInterface.h>
#include <DimensionVector.h>
#include <HipCounters.h>
#include <ProfileBasedSearcher.h>
#include <Info.h>
#include <KernelConfiguration.h>
#include <KernelResult.h>
#include <KttException.h>
#include <KttTypes.h>
#include <Searcher.h>
#include <StopCondition.h>
#include <Tuner.h>
